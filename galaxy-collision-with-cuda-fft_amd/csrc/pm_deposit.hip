// Deterministic CIC deposit: integer cell keys -> stable LSD radix sort -> segmented
// reduction of per-cell corner sums -> node assembly. Replaces the float-atomic scatter of
// reference deposit.py:47-82 `_deposit_cic`.
//
// Summation order (the contract the CPU oracle restates):
//   1. A particle's contribution to corner c of its cell is ((m*wx)*wy)*wz with
//      m = mass * (1/((dx*dx)*dx)) (the reference's `m * wx * wy * wz`).
//   2. Particles are ordered by a stable sort on the extended cell key.
//   3. The sorted array is cut into aligned chunks of kChunk positions. A cell's particles
//      that fall into one chunk form a "piece", summed sequentially in sorted order.
//   4. A cell's pieces (in order) are summed by the aligned pairwise tree: at level s,
//      element l (a multiple of 2s) absorbs element l+s when it exists; odd leftovers
//      pass up unchanged.
//   5. Node value = 0 + sum over corners c = 0..7 (bits cx,cy,cz) of the corner-c sum of
//      cell (i-cx, j-cy, k-cz), skipping absent cells, in that order.
#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>

#include "pm_cic.h"
#include "pm_common.h"
#include "pm_deposit_bricked.h"
#include "pm_deposit_fixed.h"
#include "pm_radix.h"
#include "pm_timing.h"

namespace {

constexpr int kChunk = 16;           // positions per piece chunk (one thread each)
constexpr int kSortWaves = 4;        // waves per radix workgroup
constexpr int kSortItems = 16;       // keys per lane
constexpr int kTile = kSortWaves * 64 * kSortItems;  // 4096 keys per radix tile
constexpr int kMaxDigitBits = pm::radix::kMaxDigitBits;  // 25-bit cell keys: 3 passes
static_assert(kTile == pm::radix::kTileKeys, "radix tile size");
constexpr int kSmallPieces = 8;      // open cells with <= this many pieces finish in one thread
constexpr int kWavePieces = 256;     // ... in one wave; larger ones in a workgroup
constexpr int kBigBatch = 1024;      // pieces per batch for large cells

struct BigCell {
  int64_t slot;   // sorted position of the cell's first particle
  int64_t chunk;  // chunk holding that first particle
  int64_t pieces; // number of pieces
};

// ---------------------------------------------------------------------------------------
// 1. keys (skipped when the caller supplies them, e.g. from the drift pass)
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) key_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    int64_t np, int n, float dx, uint32_t sentinel, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ keys_out) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  // all three loads first (the && below would otherwise order them)
  const float x = px[p], y = py[p], z = pz[p];
  int ix = 0, iy = 0, iz = 0;
  float fx, fy, fz;
  const bool in = pm::cic_axis(x, dx, n, ix, fx) && pm::cic_axis(y, dx, n, iy, fy) &&
                  pm::cic_axis(z, dx, n, iz, fz);
  const uint32_t key = in ? pm::cell_key(ix, iy, iz, n) : sentinel;
  keys[p] = key;
  if (keys_out) keys_out[p] = key;
}


// ---------------------------------------------------------------------------------------
// 2. stable LSD radix sort (reduce-then-scan; no inter-workgroup waiting)
// ---------------------------------------------------------------------------------------
// Per-tile digit histogram, stored digit-major: hist[d * tiles + tile].
__global__ void __launch_bounds__(256) radix_hist_kernel(const uint32_t* __restrict__ keys,
                                                         int64_t np, int tiles, int shift,
                                                         int bits, uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[1 << kMaxDigitBits];
  const int buckets = 1 << bits;
  const uint32_t mask = static_cast<uint32_t>(buckets - 1);
  for (int d = threadIdx.x; d < buckets; d += blockDim.x) cnt[d] = 0;
  __syncthreads();
  // tile number in XCD-aware order: neighbouring tiles, whose counts share cache lines of
  // the digit-major histogram, run on one XCD
  const int64_t tile = pm::xcd_block();
  const int64_t base = tile * kTile;
  // Counts do not depend on the order of the adds (plain LDS atomics). Cell-ordered input
  // is highly clustered, so each wave first collapses runs of equal digits over its 256
  // consecutive keys (lane l holds keys 4l .. 4l+3) and adds each run's length once.
  if (base + kTile <= np) {
    const uint4* q4 = reinterpret_cast<const uint4*>(keys + base);
    const int lane = threadIdx.x & 63;
    // the tile's loads all leave before the first count (one round trip per tile)
    uint4 qs[kTile / 1024];
#pragma unroll
    for (int e = 0; e < kTile / 1024; ++e) qs[e] = q4[e * 256 + threadIdx.x];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int e = 0; e < kTile / 1024; ++e) {
      const uint4 q = qs[e];
      const uint32_t d[4] = {(q.x >> shift) & mask, (q.y >> shift) & mask,
                             (q.z >> shift) & mask, (q.w >> shift) & mask};
      const uint32_t prev = __shfl_up(d[3], 1, 64);   // digit before this lane's first key
      const uint32_t next = __shfl_down(d[0], 1, 64); // digit after this lane's last key
      bool start[4], end[4];
      start[0] = lane == 0 || prev != d[0];
      end[3] = lane == 63 || next != d[3];
#pragma unroll
      for (int j = 1; j < 4; ++j) start[j] = d[j] != d[j - 1];
#pragma unroll
      for (int j = 0; j < 3; ++j) end[j] = d[j] != d[j + 1];
      int last = -1;  // last run start in this lane
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (start[j]) last = j;
      // start of the run that enters this lane: the last run start of the nearest earlier
      // lane that has one
      const uint64_t has = __ballot(last >= 0);
      const uint64_t below = has & ((1ull << lane) - 1ull);
      const int src = below ? 63 - __builtin_clzll(below) : 0;
      const int src_last = __shfl(last, src, 64);
      int run_start = 4 * src + src_last;  // valid when the lane's first key continues a run
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (start[j]) run_start = 4 * lane + j;
        if (end[j]) atomicAdd(&cnt[d[j]], static_cast<uint32_t>(4 * lane + j - run_start + 1));
      }
    }
  } else {
    for (int64_t p = base + threadIdx.x; p < np; p += blockDim.x)
      atomicAdd(&cnt[(keys[p] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < buckets; d += blockDim.x)
    hist[static_cast<int64_t>(d) * tiles + tile] = cnt[d];
}

// Block-wide exclusive scan of one value per thread (256 threads = 4 waves).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total,
                                                         uint32_t* s_wave) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t wave_off = 0;
  total = 0;
  for (int w = 0; w < 4; ++w) {
    const uint32_t s = s_wave[w];
    if (w < wave) wave_off += s;
    total += s;
  }
  __syncthreads();
  return wave_off + incl - v;
}

// One workgroup per digit: exclusive scan of hist[d][*] in place; digit_tot[d] = total.
__global__ void __launch_bounds__(256) radix_scan_kernel(uint32_t* __restrict__ hist, int tiles,
                                                         uint32_t* __restrict__ digit_tot) {
  __shared__ uint32_t s_wave[4];
  uint32_t* row = hist + static_cast<int64_t>(blockIdx.x) * tiles;
  uint32_t carry = 0;
  for (int base = 0; base < tiles; base += 256 * 4) {
    uint32_t v[4];
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = base + threadIdx.x * 4 + q;
      v[q] = t < tiles ? row[t] : 0u;
      sum += v[q];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, total, s_wave) + carry;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = base + threadIdx.x * 4 + q;
      if (t < tiles) row[t] = run;
      run += v[q];
    }
    carry += total;
  }
  if (threadIdx.x == 0) digit_tot[blockIdx.x] = carry;
}


// Stable scatter of one tile. Lane l of wave w, round r holds tile position
// w*64*kSortItems + r*64 + l, so (wave, round, lane) order is input order. A null vals_in
// stands for the identity (the first pass sorts particle indices).
__global__ void __launch_bounds__(256) radix_scatter_kernel(
    const uint32_t* __restrict__ keys_in, const int32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ digit_tot, int64_t np,
    int tiles, int shift, int bits) {
  __shared__ uint32_t wave_cnt[kSortWaves][1 << kMaxDigitBits];
  __shared__ uint32_t tile_base[1 << kMaxDigitBits];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int buckets = 1 << bits;
  const uint32_t mask = static_cast<uint32_t>(buckets - 1);
  const int64_t tile = pm::xcd_block();  // as in radix_hist_kernel
  // the tile's keys (and values) are loaded first: their latency overlaps the digit-start
  // scan below, which does not need them
  const uint64_t lanes_below = (1ull << lane) - 1ull;
  const int64_t wave_start =
      tile * kTile + static_cast<int64_t>(wave) * 64 * kSortItems;
  uint32_t k[kSortItems];
  int32_t v[kSortItems];
  uint32_t rank[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t p = wave_start + r * 64 + lane;
    const bool valid = p < np;
    k[r] = valid ? keys_in[p] : 0u;
    v[r] = valid ? (vals_in ? vals_in[p] : static_cast<int32_t>(p)) : 0;
  }
  for (int d = threadIdx.x; d < kSortWaves * (1 << kMaxDigitBits); d += blockDim.x)
    (&wave_cnt[0][0])[d] = 0;
  {
    // digit starts: exclusive scan of the digit totals (2 digits per thread)
    static_assert(2 * 256 >= (1 << kMaxDigitBits), "two digits per thread");
    __shared__ uint32_t s_wave[4];
    const int d0 = 2 * threadIdx.x;
    const uint32_t t0 = d0 < buckets ? digit_tot[d0] : 0u;
    const uint32_t t1 = d0 + 1 < buckets ? digit_tot[d0 + 1] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(t0 + t1, total, s_wave);
    if (d0 < buckets) tile_base[d0] = ex + hist[static_cast<int64_t>(d0) * tiles + tile];
    if (d0 + 1 < buckets)
      tile_base[d0 + 1] = ex + t0 + hist[static_cast<int64_t>(d0 + 1) * tiles + tile];
  }
  __syncthreads();
  // Rounds whose valid lanes all hold one digit (common for cell-ordered input) need no
  // per-bit ballots, and a run of such rounds keeps that digit's wave count in a register
  // (c_digit, c_count: wave-uniform) instead of an LDS write-then-read per round; the
  // count goes back to LDS before any other digit's count is touched.
  uint32_t c_digit = 0xFFFFFFFFu, c_count = 0;
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool valid = wave_start + r * 64 + lane < np;
    const uint32_t d = (k[r] >> shift) & mask;
    const uint64_t vmask = __ballot(valid);
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const bool uniform = __ballot(valid && d == d0) == vmask;
    const uint32_t nvalid = static_cast<uint32_t>(__popcll(vmask));
    if (uniform) {
      if (d0 != c_digit) {
        if (c_digit != 0xFFFFFFFFu) wave_cnt[wave][c_digit] = c_count;
        c_digit = d0;
        c_count = wave_cnt[wave][d0];
      }
      rank[r] = c_count + static_cast<uint32_t>(__popcll(vmask & lanes_below));
      c_count += nvalid;
    } else {
      if (c_digit != 0xFFFFFFFFu) {
        wave_cnt[wave][c_digit] = c_count;
        c_digit = 0xFFFFFFFFu;
      }
      // lanes with the same digit: one ballot per digit bit
      uint64_t match = vmask;
      for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bm = __ballot(bit);
        match &= bit ? bm : ~bm;
      }
      const uint32_t before = static_cast<uint32_t>(__popcll(match & lanes_below));
      const uint32_t base = wave_cnt[wave][d];
      rank[r] = base + before;
      if (valid && before == 0)
        wave_cnt[wave][d] = base + static_cast<uint32_t>(__popcll(match));
    }
  }
  if (c_digit != 0xFFFFFFFFu) wave_cnt[wave][c_digit] = c_count;
  __syncthreads();
  // exclusive prefix of the per-wave digit counts, in wave order
  for (int d = threadIdx.x; d < buckets; d += blockDim.x) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
      const uint32_t c = wave_cnt[w][d];
      wave_cnt[w][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t p = wave_start + r * 64 + lane;
    if (p < np) {
      const uint32_t d = (k[r] >> shift) & mask;
      const int64_t dst = static_cast<int64_t>(tile_base[d]) + wave_cnt[wave][d] + rank[r];
      keys_out[dst] = k[r];
      vals_out[dst] = v[r];
    }
  }
}

// ---------------------------------------------------------------------------------------
// 3. segmented reduction
// ---------------------------------------------------------------------------------------
// A cell's 8 corner sums are one 32-byte record, sums[slot*8 + c] (slot = sorted position
// of the cell's first particle): written with two 16-byte stores, and the 8 nodes that
// read a cell's corners share its cache line.
__device__ __forceinline__ void store_cell(float* __restrict__ sums, int64_t slot,
                                           const float v[8]) {
  float4* dst = reinterpret_cast<float4*>(sums + slot * 8);
  dst[0] = make_float4(v[0], v[1], v[2], v[3]);
  dst[1] = make_float4(v[4], v[5], v[6], v[7]);
}

// Corner terms of one particle: r = (fx, fy, fz, mass / dx^3).
__device__ __forceinline__ void corner_terms(const float4 r, float t[8]) {
  const float wx0 = 1.0f - r.x, wy0 = 1.0f - r.y, wz0 = 1.0f - r.z;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float wx = (c & 4) ? r.x : wx0;
    const float wy = (c & 2) ? r.y : wy0;
    const float wz = (c & 1) ? r.z : wz0;
    t[c] = r.w * wx * wy * wz;
  }
}

// Piece sums in two phases per workgroup of kPieceChunks chunks (kPiecePos positions).
// Phase 1, one lane per sorted position (coalesced keys and indices, gathers in near-input
// order): the particle's 8 corner terms go to LDS, chunk by chunk. Phase 2, eight lanes per
// chunk, lane c summing corner c: the chunk's runs of equal keys are summed strictly in
// order (acc = t0; acc = acc + t1; ...), the sequential piece sum of the contract above, bit
// for bit. Where a run ends its lanes write it: complete cells straight to sums/map, a
// chunk's continuation piece to head[chunk], its open last piece to tail[chunk]; lane 0 of
// the chunk writes tail_slot (sorted position of the open cell's first particle, or -1) and
// head_end (1 when the continued cell ends in this chunk), and the 8 lanes of a cell's first
// run mark the bricks of its 8 corner nodes.
constexpr int kPieceChunks = 32;                     // 8 lanes per chunk, 256 lanes
constexpr int kPiecePos = kPieceChunks * kChunk;     // 512 positions per workgroup
constexpr int kTermStride = 8 * kChunk + 8;          // floats per chunk in LDS: lanes
                                                     // (chunk q, corner c) hit bank 8q + c
static_assert(kPieceChunks * 8 == 256, "eight lanes per chunk in a 256-lane workgroup");
static_assert(kPiecePos % 256 == 0, "phase 1 covers whole positions per lane");

// Node bricks of kBrick^3 nodes: bricks[b] != 0 when some occupied cell has a corner node in
// brick b. Cell (ex, ey, ez) (extended indices, key (ex*m + ey)*m + ez, m = n+1) holds
// corner c of node (ex-1+cx, ey-1+cy, ez-1+cz); node assembly skips unmarked bricks (all
// their nodes are 0).
constexpr int kBrick = pm::kNodeBrick;

__device__ __forceinline__ void mark_corner_brick(uint32_t key, int c, int n,
                                                  const pm::FastDiv& m1,
                                                  uint8_t* __restrict__ bricks) {
  // key = (ex * m + ey) * m + ez with m = n + 1 (m1 divides by m)
  const uint32_t m = m1.d;
  const uint32_t q = pm::fast_div(key, m1);
  const int ez = static_cast<int>(key - q * m);
  const uint32_t ex_u = pm::fast_div(q, m1);
  const int ey = static_cast<int>(q - ex_u * m);
  const int ex = static_cast<int>(ex_u);
  const int nb = (n + kBrick - 1) / kBrick;
  const int i = ex - 1 + (c >> 2), j = ey - 1 + ((c >> 1) & 1), k = ez - 1 + (c & 1);
  if (i >= 0 && i < n && j >= 0 && j < n && k >= 0 && k < n)
    bricks[((i / kBrick) * nb + j / kBrick) * nb + k / kBrick] = 1;
}

__global__ void __launch_bounds__(256) piece_kernel(
    const uint32_t* __restrict__ skey, const int32_t* __restrict__ sidx,
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ mass, float dx, float inv_vol, int64_t np, uint32_t sentinel,
    int64_t chunks, float* __restrict__ sums, int32_t* __restrict__ map,
    float* __restrict__ head, float* __restrict__ tail, int64_t* __restrict__ tail_slot,
    uint8_t* __restrict__ head_end, int n, uint8_t* __restrict__ bricks, pm::FastDiv m1) {
  __shared__ float4 terms4[kPieceChunks * kTermStride / 4];
  __shared__ uint32_t keys_l[kPiecePos + 2];  // keys of positions base-1 .. base+kPiecePos
  float* terms = reinterpret_cast<float*>(terms4);
  const int64_t blk = pm::xcd_block();
  const int64_t base = blk * kPiecePos;
  const int tid = threadIdx.x;

  // phase 1: corner terms, one lane per position (positions >= np get key ~0u, terms 0).
  // All loads of a lane's positions are issued before any is used: keys and indices, then
  // the particle gathers (any index is valid, so they need not wait for the key test).
  constexpr int kPer = kPiecePos / 256;
  uint32_t key[kPer];
  int32_t q[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int64_t g = base + r * 256 + tid;
    key[r] = g < np ? skey[g] : 0xFFFFFFFFu;
    q[r] = g < np ? sidx[g] : 0;
  }
  float x[kPer], y[kPer], z[kPer], w[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    x[r] = px[q[r]];
    y[r] = py[q[r]];
    z[r] = pz[q[r]];
    w[r] = mass[q[r]];
  }
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int l = r * 256 + tid;
    float t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (key[r] < sentinel) {
      // the particle's offsets in its cell, recomputed from its coordinates exactly as its
      // key was (a live key means all three axes are in range)
      int ix, iy, iz;
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
      (void)pm::cic_axis(x[r], dx, n, ix, fx);
      (void)pm::cic_axis(y[r], dx, n, iy, fy);
      (void)pm::cic_axis(z[r], dx, n, iz, fz);
      corner_terms(make_float4(fx, fy, fz, w[r] * inv_vol), t);
    }
    keys_l[l + 1] = key[r];
    float4* dst = terms4 + ((l / kChunk) * kTermStride + (l % kChunk) * 8) / 4;
    dst[0] = make_float4(t[0], t[1], t[2], t[3]);
    dst[1] = make_float4(t[4], t[5], t[6], t[7]);
  }
  if (tid == 0) keys_l[0] = base > 0 ? skey[base - 1] : 0xFFFFFFFFu;
  if (tid == 1) keys_l[kPiecePos + 1] = base + kPiecePos < np ? skey[base + kPiecePos] : 0xFFFFFFFFu;
  __syncthreads();

  // phase 2: lane c of chunk q sums corner c over the chunk's runs, in order
  const int lq = tid >> 3;
  const int c = tid & 7;
  const int64_t ch = blk * kPieceChunks + lq;
  if (ch >= chunks) return;  // no barrier below
  const int64_t g0 = ch * kChunk;
  const uint32_t* kk = keys_l + 1 + lq * kChunk;  // kk[-1]: key before the chunk
  const float* tt = terms + lq * kTermStride + c;
  const uint32_t prevk = kk[-1];
  float acc = 0.0f;
  uint32_t rk = 0u;
  int rs = 0;
  bool open = false;
  bool closing = false;  // the run continued from the previous chunk ends here
  // ends the open run (rs, rk) with sum acc
  auto flush = [&](bool ends) {
    const bool starts = rs > 0 || prevk != rk;
    if (starts && ends) {
      sums[(g0 + rs) * 8 + c] = acc;
      if (c == 0) map[rk] = static_cast<int32_t>(g0 + rs);
    } else if (starts) {
      tail[ch * 8 + c] = acc;
    } else {
      head[ch * 8 + c] = acc;
    }
    if (starts) mark_corner_brick(rk, c, n, m1, bricks);
    if (!starts && ends) closing = true;
  };
  uint32_t kr[kChunk];
  float tr[kChunk];
#pragma unroll
  for (int j = 0; j < kChunk; ++j) {
    kr[j] = kk[j];
    tr[j] = tt[j * 8];
  }
#pragma unroll
  for (int j = 0; j < kChunk; ++j) {
    const uint32_t key = kr[j];
    const float t = tr[j];
    const bool live = key < sentinel;
    const bool same = open && live && key == rk;
    if (open && !same) flush(true);
    acc = same ? acc + t : t;
    if (!same) {
      rk = key;
      rs = j;
    }
    open = live;
  }
  // a run still open reaches the chunk's last position
  const bool last_ends = kk[kChunk] != rk;
  if (open) flush(last_ends);
  if (c == 0) {
    tail_slot[ch] = (open && (rs > 0 || prevk != rk) && !last_ends) ? g0 + rs : -1;
    head_end[ch] = closing ? 1 : 0;
  }
}

// First sorted position after `from` whose key differs from `key` (keys are sorted).
// Galloping: probe from+1, +2, +4, ... until a different key (or the end), then bisect the
// last gap, so a cell of L particles costs about 2 log2(L) mostly cached loads.
__device__ __forceinline__ int64_t cell_end(const uint32_t* __restrict__ skey, int64_t from,
                                            int64_t np, uint32_t key) {
  int64_t lo = from, step = 1, hi = np;  // invariant: skey[lo] == key, answer in (lo, hi]
  while (lo + step < np) {
    if (skey[lo + step] != key) {
      hi = lo + step;
      break;
    }
    lo += step;
    step <<= 1;
  }
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (skey[mid] == key) lo = mid; else hi = mid;
  }
  return hi;
}

// Cells whose pieces span chunks: small ones (<= kSmallPieces pieces) finish here with an
// unrolled aligned pairwise tree; up to kWavePieces pieces are marked in mid_pieces[chunk]
// for mid_cell_kernel (no shared counter: 10^4 such cells would serialize on one
// device-scope atomic); larger ones are listed for big_batch_kernel.
__global__ void __launch_bounds__(256) join_kernel(
    const uint32_t* __restrict__ skey, int64_t np, int64_t chunks,
    const int64_t* __restrict__ tail_slot, const uint8_t* __restrict__ head_end,
    const float* __restrict__ head,
    const float* __restrict__ tail, float* __restrict__ sums, int64_t cap,
    int32_t* __restrict__ map, BigCell* __restrict__ big, uint32_t* __restrict__ big_count,
    int32_t* __restrict__ mid_pieces) {
  const int64_t ch = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (ch >= chunks) return;
  const int64_t slot = tail_slot[ch];
  mid_pieces[ch] = 0;
  if (slot < 0) return;
  const uint32_t key = skey[slot];
  // the cell's chunks ch+1 .. last continue it; last is the first with head_end set, read
  // as 8 bytes (head_end carries 16 bytes of padding)
  const uint64_t base = static_cast<uint64_t>(ch) + 1;
  const uint64_t* words = reinterpret_cast<const uint64_t*>(head_end);
  const unsigned sh = static_cast<unsigned>(base & 7u) * 8u;
  const uint64_t w0 = words[base >> 3];
  const uint64_t win = sh == 0 ? w0 : (w0 >> sh) | (words[(base >> 3) + 1] << (64u - sh));
  int64_t pieces;
  if (win != 0 && __builtin_ctzll(win) / 8 + 2 <= kSmallPieces) {
    pieces = __builtin_ctzll(win) / 8 + 2;
  } else {
    const int64_t end = cell_end(skey, slot, np, key);
    pieces = (end - 1) / kChunk - ch + 1;
  }
  if (pieces > kWavePieces) {  // rare: listed for big_batch_kernel
    const uint32_t at = atomicAdd(big_count, 1u);
    big[at] = BigCell{slot, ch, pieces};
    return;
  }
  if (pieces > kSmallPieces) {  // found by mid_cell_kernel's scan
    mid_pieces[ch] = static_cast<int32_t>(pieces);
    return;
  }
  float p[kSmallPieces][8];
#pragma unroll
  for (int c = 0; c < 8; ++c) p[0][c] = tail[ch * 8 + c];
#pragma unroll
  for (int q = 1; q < kSmallPieces; ++q) {
    if (q < pieces) {
#pragma unroll
      for (int c = 0; c < 8; ++c) p[q][c] = head[(ch + q) * 8 + c];
    }
  }
#pragma unroll
  for (int s = 1; s < kSmallPieces; s <<= 1) {
#pragma unroll
    for (int l = 0; l + s < kSmallPieces; l += 2 * s) {
      if (l + s < pieces) {
#pragma unroll
        for (int c = 0; c < 8; ++c) p[l][c] = p[l][c] + p[l + s][c];
      }
    }
  }
  store_cell(sums, slot, p[0]);
  map[key] = static_cast<int32_t>(slot);
}

// Piece q of a cell whose first piece closes chunk `chunk` (q = 0: that tail piece).
__device__ __forceinline__ void load_piece(const float* __restrict__ head,
                                           const float* __restrict__ tail, int64_t chunk,
                                           int64_t q, float v[8]) {
  const float* src = q == 0 ? tail + chunk * 8 : head + (chunk + q) * 8;
  const float4 a = *reinterpret_cast<const float4*>(src);
  const float4 b = *reinterpret_cast<const float4*>(src + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

// Cells of kSmallPieces+1 .. kWavePieces pieces: one wave each, the first stage of
// big_batch_kernel's batch (4 pieces per lane, aligned shuffle tree), whose result for a
// batch of at most 256 pieces is the wave-0 value, so the sums are bitwise the same.
__global__ void __launch_bounds__(256) mid_cell_kernel(
    const uint32_t* __restrict__ skey, const int64_t* __restrict__ tail_slot,
    const int32_t* __restrict__ mid_pieces, int64_t chunks, const float* __restrict__ head,
    const float* __restrict__ tail, float* __restrict__ sums, int32_t* __restrict__ map) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  // each workgroup scans windows of 64 chunks; its waves take the window's marked chunks in
  // turn (a window in a dense region starts up to 7 such cells: shared by the waves they do
  // not form one wave's serial chain)
  const int64_t windows = (chunks + 63) / 64;
  for (int64_t win = blockIdx.x; win < windows; win += gridDim.x) {
    const int64_t base = win * 64;
    const int32_t mine = base + lane < chunks ? mid_pieces[base + lane] : 0;
    uint64_t todo = __ballot(mine > 0);
    int rank = 0;
    while (todo) {
      const int l = __builtin_ctzll(todo);
      todo &= todo - 1;
      if ((rank++ % nwaves) != wid) continue;
      const int64_t chunk = base + l;
      const int64_t m = __shfl(mine, l, 64);
      const int64_t slot = tail_slot[chunk];
      const int64_t q0 = 4 * lane;
      float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (q0 < m) {
        float w[8];
        load_piece(head, tail, chunk, q0, v);
        if (q0 + 1 < m) {
          load_piece(head, tail, chunk, q0 + 1, w);
#pragma unroll
          for (int c = 0; c < 8; ++c) v[c] = v[c] + w[c];
        }
        if (q0 + 2 < m) {
          float u[8];
          load_piece(head, tail, chunk, q0 + 2, w);
          if (q0 + 3 < m) {
            load_piece(head, tail, chunk, q0 + 3, u);
#pragma unroll
            for (int c = 0; c < 8; ++c) w[c] = w[c] + u[c];
          }
#pragma unroll
          for (int c = 0; c < 8; ++c) v[c] = v[c] + w[c];
        }
      }
#pragma unroll
      for (int sft = 1; sft < 64; sft <<= 1) {
        const bool take = (lane % (2 * sft)) == 0 && (q0 + 4 * sft) < m;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const float o = __shfl_down(v[c], sft, 64);
          if (take) v[c] = v[c] + o;
        }
      }
      // lane 0 holds the 8 corner sums
      if (lane == 0) {
        store_cell(sums, slot, v);
        map[skey[slot]] = static_cast<int32_t>(slot);
      }
    }
  }
}

// One workgroup (4 waves) per cell with more than kWavePieces pieces. Batches of
// kBigBatch = 1024 pieces, aligned to the cell's first piece: each lane reduces 4
// consecutive pieces, the wave continues the aligned pairwise tree with shuffles (lane l
// absorbs lane l+s when that lane holds pieces), and the 4 wave results are combined the
// same way. Batch results are merged by a binary-counter cascade (push; merge while the two
// top partials cover equal piece counts) and a final right-nested combine, which is the
// same aligned tree one level up.
// Cells of more than kWavePieces pieces, in two kernels so that a few huge cells (the cores
// of a galaxy at 100M particles hold tens of thousands of pieces each) spread over the
// whole device. big_batch_kernel sums each aligned batch of kBigBatch pieces of every such
// cell (4 pieces per lane, shuffle levels, the wave combine: the levels of the aligned
// pairwise tree inside the batch), one workgroup per batch, and stores the batch sum in the
// batch's first piece slot (tail[chunk] for batch 0, head[chunk + kBigBatch*b] after, read
// by no one else). big_merge_kernel then folds a cell's batch sums in order with the binary
// counter of the tree's upper levels (equal sizes merge, sizes double) and the final
// right-to-left fold: the same float operations in the same order as one workgroup walking
// the batches.
__device__ __forceinline__ float* batch_slot(float* __restrict__ head, float* __restrict__ tail,
                                             int64_t chunk, int64_t b) {
  return b == 0 ? tail + chunk * 8 : head + (chunk + b * kBigBatch) * 8;
}

// big_off[i] = batches of the listed cells before cell i; big_off[count] = all batches.
__global__ void __launch_bounds__(256) big_scan_kernel(const BigCell* __restrict__ big,
                                                       const uint32_t* __restrict__ big_count,
                                                       int64_t* __restrict__ big_off) {
  __shared__ uint32_t s_wave[4];
  const uint32_t count = *big_count;
  int64_t running = 0;
  for (uint32_t base = 0; base < count; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t nb =
        i < count ? static_cast<uint32_t>((big[i].pieces + kBigBatch - 1) / kBigBatch) : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(nb, total, s_wave);
    if (i < count) big_off[i] = running + ex;
    running += total;
  }
  if (threadIdx.x == 0) big_off[count] = running;
}

__global__ void __launch_bounds__(256) big_batch_kernel(
    const BigCell* __restrict__ big, const uint32_t* __restrict__ big_count,
    const int64_t* __restrict__ big_off, float* __restrict__ head, float* __restrict__ tail) {
  __shared__ float s_wave[4][8];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t count = *big_count;
  if (count == 0) return;
  const int64_t total = big_off[count];
  for (int64_t f = blockIdx.x; f < total; f += gridDim.x) {
    // the cell of batch f: the last i with big_off[i] <= f (offsets increase: nb >= 1)
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (big_off[mid] <= f) lo = mid; else hi = mid - 1;
    }
    const BigCell cell = big[lo];
    const int64_t b = f - big_off[lo];
    {
      const int64_t base = b * kBigBatch;
      const int64_t m = cell.pieces - base < kBigBatch ? cell.pieces - base : kBigBatch;
      // lane holds pieces base + 256*wave + 4*lane + {0..3}
      const int64_t q0 = 256 * wave + 4 * lane;  // relative to base
      float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (q0 < m) {
        float w[8];
        load_piece(head, tail, cell.chunk, base + q0, v);
        if (q0 + 1 < m) {
          load_piece(head, tail, cell.chunk, base + q0 + 1, w);
#pragma unroll
          for (int c = 0; c < 8; ++c) v[c] = v[c] + w[c];
        }
        if (q0 + 2 < m) {
          float u[8];
          load_piece(head, tail, cell.chunk, base + q0 + 2, w);
          if (q0 + 3 < m) {
            load_piece(head, tail, cell.chunk, base + q0 + 3, u);
#pragma unroll
            for (int c = 0; c < 8; ++c) w[c] = w[c] + u[c];
          }
#pragma unroll
          for (int c = 0; c < 8; ++c) v[c] = v[c] + w[c];
        }
      }
#pragma unroll
      for (int sft = 1; sft < 64; sft <<= 1) {
        const bool take = (lane % (2 * sft)) == 0 && (q0 + 4 * sft) < m;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const float o = __shfl_down(v[c], sft, 64);
          if (take) v[c] = v[c] + o;
        }
      }
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 8; ++c) s_wave[wave][c] = v[c];
      }
      __syncthreads();  // every piece of the batch has been loaded
      if (tid < 8) {
        const int c = tid;
        float w0 = s_wave[0][c], w2 = s_wave[2][c];
        if (256 < m) w0 = w0 + s_wave[1][c];
        if (768 < m) w2 = w2 + s_wave[3][c];
        batch_slot(head, tail, cell.chunk, b)[c] = 512 < m ? w0 + w2 : w0;
      }
      __syncthreads();  // s_wave is reused by the next batch
    }
  }
}

// Eight lanes per cell, lane c folding corner c.
constexpr int kMergeDepth = 40;
__global__ void __launch_bounds__(256) big_merge_kernel(
    const uint32_t* __restrict__ skey, const BigCell* __restrict__ big,
    const uint32_t* __restrict__ big_count, float* __restrict__ head,
    float* __restrict__ tail, float* __restrict__ sums, int32_t* __restrict__ map) {
  __shared__ float stack[kMergeDepth][256];
  __shared__ int64_t stack_size[kMergeDepth][32];
  const int tid = threadIdx.x;
  const int c = tid & 7;
  const int g = tid >> 3;  // cell group of this workgroup
  const uint32_t count = *big_count;
  for (uint32_t bcell = blockIdx.x * 32 + g; bcell < count; bcell += gridDim.x * 32) {
    const BigCell cell = big[bcell];
    const int64_t nb = (cell.pieces + kBigBatch - 1) / kBigBatch;
    int depth = 0;
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t m = cell.pieces - b * kBigBatch < kBigBatch ? cell.pieces - b * kBigBatch
                                                                : kBigBatch;
      float val = batch_slot(head, tail, cell.chunk, b)[c];
      int d = depth;
      int64_t sz = m;
      while (d > 0 && stack_size[d - 1][g] == sz) {
        val = stack[d - 1][tid] + val;
        sz *= 2;
        --d;
      }
      stack[d][tid] = val;
      // the eight lanes of the group read stack_size above and write the same value here
      stack_size[d][g] = sz;
      depth = d + 1;
    }
    float acc = stack[depth - 1][tid];
    for (int d = depth - 2; d >= 0; --d) acc = stack[d][tid] + acc;
    sums[cell.slot * 8 + c] = acc;
    if (c == 0) map[skey[cell.slot]] = static_cast<int32_t>(cell.slot);
  }
}

// ---------------------------------------------------------------------------------------
// 4. node assembly
// ---------------------------------------------------------------------------------------
// rho = 0 with 16-byte stores where n^3 allows (the nodes of unmarked bricks stay 0).
__global__ void __launch_bounds__(256) zero_rho_kernel(float* __restrict__ rho, int64_t count,
                                                       bool vec) {
  const int64_t c0 = 4 * (pm::xcd_block() * blockDim.x + threadIdx.x);
  if (c0 >= count) return;
  if (vec && c0 + 3 < count) {
    *reinterpret_cast<float4*>(rho + c0) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  for (int64_t c = c0; c < count && c < c0 + 4; ++c) rho[c] = 0.0f;
}

// The marked bricks, compacted into list[0..*count) (any order: each is processed on its
// own). One workgroup of 1024 lanes; each lane tests 32 consecutive marks, read as two
// 16-byte loads issued before any is used.
constexpr int kListLanes = 1024;
constexpr int kListSpan = 32;  // marks per lane and round
__global__ void __launch_bounds__(kListLanes) list_bricks_kernel(const uint8_t* __restrict__ bricks,
                                                                 int64_t nbricks,
                                                                 int32_t* __restrict__ list,
                                                                 uint32_t* __restrict__ count) {
  __shared__ uint32_t s_count;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  if (tid == 0) s_count = 0u;
  __syncthreads();
  for (int64_t base = 0; base < nbricks; base += int64_t(kListLanes) * kListSpan) {
    const int64_t lo = base + int64_t(tid) * kListSpan;
    uint8_t mk[kListSpan];
    if (lo + kListSpan <= nbricks) {
      const uint4 a = reinterpret_cast<const uint4*>(bricks + lo)[0];
      const uint4 b = reinterpret_cast<const uint4*>(bricks + lo)[1];
      const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int q = 0; q < kListSpan; ++q) mk[q] = (w[q >> 2] >> (8 * (q & 3))) & 0xFFu;
    } else {
#pragma unroll
      for (int q = 0; q < kListSpan; ++q) mk[q] = lo + q < nbricks ? bricks[lo + q] : 0;
    }
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < kListSpan; ++q) c += mk[q] != 0;
    // wave prefix of the counts, one LDS atomic per wave
    uint32_t incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    const uint32_t wave_total = __shfl(incl, 63, 64);
    uint32_t wave_base = 0;
    if (lane == 0 && wave_total) wave_base = atomicAdd(&s_count, wave_total);
    uint32_t pos = __shfl(wave_base, 0, 64) + incl - c;
#pragma unroll
    for (int q = 0; q < kListSpan; ++q)
      if (mk[q]) list[pos++] = static_cast<int32_t>(lo + q);
  }
  __syncthreads();
  if (tid == 0) *count = s_count;
}

// After zero_rho_kernel, the marked bricks (list_bricks_kernel), a workgroup at a time: the
// brick's 9^3 extended cells (those holding a corner of one of its nodes) are staged in
// LDS: their map entries, coalesced, and the 32-byte corner records of the occupied ones,
// each loaded once. Then each thread assembles 2 nodes from LDS. Per node the order is that
// of step 5 above: 0 + corner c = 0..7 of cell (i+1-cx, j+1-cy, k+1-cz) (extended indices),
// absent cells skipped. Unmarked bricks (some 90% at 256^3 / 10M particles) cost nothing
// beyond the zero fill.
constexpr unsigned kPersistentBlocks = 2048;  // workgroups looping over listed bricks
constexpr int kCellSpan = kBrick + 1;
constexpr int kBrickCells = kCellSpan * kCellSpan * kCellSpan;  // 729
static_assert(kBrick == 8, "two nodes per lane of a 256-lane workgroup cover an 8^3 brick");

__global__ void __launch_bounds__(256) assemble_kernel(const int32_t* __restrict__ map,
                                                       const float* __restrict__ sums, int n,
                                                       float* __restrict__ rho, bool vec,
                                                       const int32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ count) {
  __shared__ int32_t s_slot[kBrickCells];
  __shared__ float4 s_sum[kBrickCells][2];
  const int nb = (n + kBrick - 1) / kBrick;
  const int64_t m = n + 1;
  const int tid = threadIdx.x;
  // lane: row (a, bb) of the brick and nodes 2h, 2h + 1 of the row
  const int a = tid >> 5, bb = (tid >> 2) & 7, h = tid & 3;
  const uint32_t cnt = *count;
  for (uint32_t e = blockIdx.x; e < cnt; e += gridDim.x) {
    const int64_t b = list[e];
    // brick index < nb^3 <= 2^21: 32-bit division (the 64-bit sequence is several times
    // longer, and every wave runs it per brick)
    const uint32_t ub = static_cast<uint32_t>(b), unb = static_cast<uint32_t>(nb);
    const int bk = static_cast<int>(ub % unb), bj = static_cast<int>((ub / unb) % unb),
              bi = static_cast<int>(ub / (unb * unb));
    const int i0 = bi * kBrick, j0 = bj * kBrick, k0 = bk * kBrick;
    for (int q = tid; q < kBrickCells; q += blockDim.x) {
      const int ex = i0 + q / (kCellSpan * kCellSpan), ey = j0 + (q / kCellSpan) % kCellSpan,
                ez = k0 + q % kCellSpan;
      int32_t slot = -1;
      if (ex <= n && ey <= n && ez <= n) slot = map[(ex * m + ey) * m + ez];
      s_slot[q] = slot;
      if (slot >= 0) {
        const float4* src = reinterpret_cast<const float4*>(sums + static_cast<int64_t>(slot) * 8);
        s_sum[q][0] = src[0];
        s_sum[q][1] = src[1];
      }
    }
    __syncthreads();
    const int i = i0 + a, j = j0 + bb, k = k0 + 2 * h;
    if (i < n && j < n && k < n) {
      float out[2];
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int q = ((a + 1 - (c >> 2)) * kCellSpan + (bb + 1 - ((c >> 1) & 1))) * kCellSpan +
                        (2 * h + o + 1 - (c & 1));
          const int32_t slot = s_slot[q];
          const float4 half = s_sum[q][c >> 2];
          const float v = (c & 3) == 0 ? half.x : (c & 3) == 1 ? half.y : (c & 3) == 2 ? half.z
                                                                                      : half.w;
          acc = slot >= 0 ? acc + v : acc;
        }
        out[o] = acc;
      }
      float* dst = rho + (static_cast<int64_t>(i) * n + j) * n + k;
      if (vec && k + 1 < n) {  // n even and rho 8-byte aligned
        *reinterpret_cast<float2*>(dst) = make_float2(out[0], out[1]);
      } else {
        dst[0] = out[0];
        if (k + 1 < n) dst[1] = out[1];
      }
    }
    __syncthreads();  // the next brick reuses the LDS
  }
}

// Sparse reset over the bricks listed at the last deposit: the map holds entries only for
// the cells occupied then, and each of those has a corner node in a brick marked (and
// listed) then, so clearing the cells (ex, ey, ez) in [8b, 8b+8]^3 of every listed brick b
// restores map = -1 everywhere; the marks are cleared for this deposit.
__global__ void __launch_bounds__(256) reset_listed_kernel(int32_t* __restrict__ map, int n,
                                                           uint8_t* __restrict__ bricks,
                                                           const int32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ count,
                                                           uint32_t* __restrict__ big_count) {
  const int nb = (n + kBrick - 1) / kBrick;
  if (blockIdx.x == 0 && threadIdx.x == 0) *big_count = 0u;
  const int64_t m = n + 1;
  constexpr int kSpan = kBrick + 1;  // cells ex in [8bi, 8bi+8]
  const uint32_t cnt = *count;
  for (uint32_t e = blockIdx.x; e < cnt; e += gridDim.x) {
    const int64_t b = list[e];
    if (threadIdx.x == 0) bricks[b] = 0;
    // brick index < nb^3 <= 2^21: 32-bit division (the 64-bit sequence is several times
    // longer, and every wave runs it per brick)
    const uint32_t ub = static_cast<uint32_t>(b), unb = static_cast<uint32_t>(nb);
    const int bk = static_cast<int>(ub % unb), bj = static_cast<int>((ub / unb) % unb),
              bi = static_cast<int>(ub / (unb * unb));
    for (int q = threadIdx.x; q < kSpan * kSpan * kSpan; q += blockDim.x) {
      const int ex = bi * kBrick + q / (kSpan * kSpan), ey = bj * kBrick + (q / kSpan) % kSpan,
                ez = bk * kBrick + q % kSpan;
      if (ex <= n && ey <= n && ez <= n) map[(ex * m + ey) * m + ez] = -1;
    }
  }
}

// map = -1 for every extended cell, bricks = 0, no big cells listed yet.
__global__ void __launch_bounds__(256) reset_map_kernel(int32_t* __restrict__ map, int64_t cells,
                                                        uint8_t* __restrict__ bricks,
                                                        int64_t nbricks,
                                                        uint32_t* __restrict__ big_count) {
  const int64_t i = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (i < cells) map[i] = -1;
  if (i < nbricks) bricks[i] = 0;
  if (i == 0) *big_count = 0u;
}


}  // namespace

namespace pm {
namespace radix {

int plan_digits(uint64_t limit, int bits[4]) {
  int key_bits = 1;
  while ((uint64_t(1) << key_bits) <= limit) ++key_bits;
  const int passes = (key_bits + kMaxDigitBits - 1) / kMaxDigitBits;
  for (int s = 0; s < 4; ++s)
    bits[s] = s < passes ? key_bits / passes + (s < key_bits % passes ? 1 : 0) : 0;
  return passes;
}

pm_status sort(const uint32_t* keys_in, uint32_t* const keys[2], int32_t* const vals[2],
               uint32_t* hist, uint32_t* digit, int64_t np, int shift, const int* bits,
               int passes, int* out, hipStream_t stream) {
  PM_REQUIRE(passes >= 1 && passes <= 4, "radix sort: 1 to 4 passes");
  const int tiles = static_cast<int>(tiles_of(np));
  int src = 0;
  for (int s = 0; s < passes; ++s) {
    const int b = bits[s];
    PM_REQUIRE(b >= 1 && b <= kMaxDigitBits, "radix sort: digit width out of range");
    const int buckets = 1 << b;
    const uint32_t* kin = s == 0 ? keys_in : keys[src];
    {
      PM_TIMED("deposit.radix_hist", 4.0 * np, stream);
      radix_hist_kernel<<<tiles, 256, 0, stream>>>(kin, np, tiles, shift, b, hist);
      PM_CHECK_LAUNCH("radix_hist_kernel");
    }
    {
      PM_TIMED("deposit.radix_scan", 8.0 * buckets * tiles, stream);
      radix_scan_kernel<<<buckets, 256, 0, stream>>>(hist, tiles, digit);
      PM_CHECK_LAUNCH("radix_scan_kernel");
    }
    {
      // read key (+ index after the first pass), write key + index
      PM_LAUNCH_TIMED("deposit.radix_scatter", (s == 0 ? 12.0 : 16.0) * np,
                      radix_scatter_kernel, tiles, 256, 0, stream, kin,
                      s == 0 ? nullptr : vals[src], keys[1 - src], vals[1 - src], hist, digit,
                      np, tiles, shift, b);
      PM_CHECK_LAUNCH("radix_scatter_kernel");
    }
    src = 1 - src;
    shift += b;
  }
  *out = src;
  return PM_OK;
}

}  // namespace radix
}  // namespace pm

struct pm_deposit_plan {
  int n = 0;
  int64_t cap = 0;
  uint32_t sentinel = 0;
  int passes = 0;
  int digit_bits[4] = {0, 0, 0, 0};
  uint32_t* keys[2] = {nullptr, nullptr};
  int32_t* vals[2] = {nullptr, nullptr};
  uint32_t* hist = nullptr;
  uint32_t* digit = nullptr;
  float* sums = nullptr;
  float* head = nullptr;
  float* tail = nullptr;
  int64_t* tail_slot = nullptr;
  uint8_t* head_end = nullptr;  // chunks + 16 bytes (join reads 16-byte windows)
  int32_t* mid_pieces = nullptr;
  uint8_t* bricks = nullptr;    // ceil(n/8)^3 node-brick marks
  int32_t* brick_list = nullptr;   // the marked bricks of the last deposit (sorted mode)
  uint32_t* brick_count = nullptr;
  bool bricks_valid = false;    // marks of the last deposit (sorted mode only)
  const int32_t* order = nullptr;  // particle order of the last sorted-mode deposit
  bool map_clean = false;       // map is -1 outside the cells around the marked bricks
  int32_t* map = nullptr;
  BigCell* big = nullptr;
  uint32_t* big_count = nullptr;
  int64_t* big_off = nullptr;   // batch offsets of the listed big cells (chunks + 1)
  bool timing = false;
  hipEvent_t ev[5] = {};
  float last_ms[4] = {0, 0, 0, 0};
  int mode = PM_DEPOSIT_SORTED;
  unsigned long long* acc64 = nullptr;  // PM_DEPOSIT_FIXED: n^3 accumulators (kept zeroed)
  unsigned int* maxbits = nullptr;
  pm::depbrick::State* brick = nullptr;  // PM_DEPOSIT_BRICKED
};

namespace {

void free_plan(pm_deposit_plan* p) {
  if (!p) return;
  void* bufs[] = {p->keys[0], p->keys[1], p->vals[0], p->vals[1], p->hist, p->digit,
                  p->sums, p->head, p->tail, p->tail_slot, p->head_end, p->mid_pieces, p->bricks, p->brick_list, p->brick_count, p->map, p->big,
                  p->big_count, p->big_off, p->acc64, p->maxbits};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto& e : p->ev)
    if (e) (void)hipEventDestroy(e);
  pm::depbrick::destroy(p->brick);
  delete p;
}

int64_t nbricks_of(int n) {
  const int64_t nb = (n + kBrick - 1) / kBrick;
  return nb * nb * nb;
}

template <typename T>
pm_status alloc(T** ptr, int64_t count, const char* what) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), sizeof(T) * (count > 0 ? count : 1));
  if (e != hipSuccess) return pm::hip_fail(what, e);
  return PM_OK;
}

// Grid and sort parameters of a plan for `cap` particles on an n^3 grid.
void init_plan_shape(pm_deposit_plan* p, int n, int64_t cap) {
  p->n = n;
  p->cap = cap;
  const int64_t m = static_cast<int64_t>(n) + 1;
  p->sentinel = static_cast<uint32_t>(m * m * m);
  p->passes = pm::radix::plan_digits(p->sentinel, p->digit_bits);
}

// Calls f(buffer pointer, element count, name) for every sorted-mode scratch buffer of a
// plan whose shape init_plan_shape set, stopping at the first error. The allocating plan
// and the no-allocation entry point lay out this one list.
template <class F>
pm_status for_each_buffer(pm_deposit_plan* p, F&& f) {
  const int64_t cap = p->cap;
  const int64_t m = static_cast<int64_t>(p->n) + 1;
  const int64_t tiles = (cap + kTile - 1) / kTile;
  const int64_t chunks = (cap + kChunk - 1) / kChunk;
  const int64_t nbricks = nbricks_of(p->n);
  pm_status st = PM_OK;
  if ((st = f(p->keys[0], cap, "hipMalloc keys")) ||
      (st = f(p->keys[1], cap, "hipMalloc keys")) ||
      (st = f(p->vals[0], cap, "hipMalloc vals")) ||
      (st = f(p->vals[1], cap, "hipMalloc vals")) ||
      (st = f(p->hist, tiles << kMaxDigitBits, "hipMalloc hist")) ||
      (st = f(p->digit, int64_t(1) << kMaxDigitBits, "hipMalloc digit")) ||
      (st = f(p->sums, 8 * cap, "hipMalloc sums")) ||
      (st = f(p->head, 8 * chunks, "hipMalloc head")) ||
      (st = f(p->tail, 8 * chunks, "hipMalloc tail")) ||
      (st = f(p->tail_slot, chunks, "hipMalloc tail_slot")) ||
      (st = f(p->head_end, chunks + 16, "hipMalloc head_end")) ||
      (st = f(p->mid_pieces, chunks, "hipMalloc mid_pieces")) ||
      (st = f(p->bricks, nbricks, "hipMalloc bricks")) ||
      (st = f(p->brick_list, nbricks, "hipMalloc brick_list")) ||
      (st = f(p->brick_count, 1, "hipMalloc brick_count")) ||
      (st = f(p->map, m * m * m, "hipMalloc map")) ||
      (st = f(p->big, chunks, "hipMalloc big")) ||
      (st = f(p->big_count, 1, "hipMalloc big_count")) ||
      (st = f(p->big_off, chunks + 1, "hipMalloc big_off")))
    return st;
  return PM_OK;
}

}  // namespace

extern "C" {

pm_status pm_deposit_plan_create(int32_t n, int64_t cap, pm_deposit_plan** out) {
  PM_REQUIRE(out, "pm_deposit_plan_create: null output");
  *out = nullptr;
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_deposit_plan_create: grid_size must be in [1, 1024]");
  PM_REQUIRE(cap >= 1 && cap < (int64_t(1) << 31),
             "pm_deposit_plan_create: max_particles must be in [1, 2^31)");
  auto* p = new pm_deposit_plan();
  init_plan_shape(p, n, cap);
  const pm_status st = for_each_buffer(p, [](auto*& ptr, int64_t count, const char* what) {
    return alloc(&ptr, count, what);
  });
  if (st != PM_OK) {
    free_plan(p);
    return st;
  }
  for (auto& e : p->ev) {
    hipError_t he = hipEventCreate(&e);
    if (he != hipSuccess) {
      free_plan(p);
      return pm::hip_fail("hipEventCreate", he);
    }
  }
  *out = p;
  return PM_OK;
}

pm_status pm_deposit_plan_destroy(pm_deposit_plan* plan) {
  free_plan(plan);
  return PM_OK;
}

pm_status pm_deposit_plan_set_mode(pm_deposit_plan* plan, int32_t mode) {
  pm::clear_error();
  PM_REQUIRE(plan, "pm_deposit_plan_set_mode: null plan");
  PM_REQUIRE(mode == PM_DEPOSIT_SORTED || mode == PM_DEPOSIT_FIXED || mode == PM_DEPOSIT_BRICKED,
             "pm_deposit_plan_set_mode: unknown mode");
  if (mode == PM_DEPOSIT_BRICKED && !plan->brick) {
    pm_status st = pm::depbrick::create(plan->n, plan->cap, &plan->brick);
    if (st != PM_OK) return st;
  }
  // a sorted-mode deposit overwrites the sort buffers that hold the binning's order
  if (plan->brick && mode != plan->mode) pm::depbrick::invalidate(plan->brick);
  if (mode == PM_DEPOSIT_FIXED && !plan->acc64) {
    const int64_t nodes = static_cast<int64_t>(plan->n) * plan->n * plan->n;
    pm_status st;
    if ((st = alloc(&plan->acc64, nodes, "hipMalloc (deposit accumulators)")) ||
        (st = alloc(&plan->maxbits, 1 + pm::depfix::kMaxBlocks, "hipMalloc (deposit scale)")))
      return st;
    if ((st = pm::depfix::zero(plan->acc64, nodes, nullptr))) return st;
    PM_TRY_HIP(hipStreamSynchronize(nullptr));
  }
  plan->mode = mode;
  return PM_OK;
}

pm_status pm_deposit_plan_brick_list(pm_deposit_plan* plan, const int32_t** list,
                                     const uint32_t** count, int64_t* max_count) {
  PM_REQUIRE(plan && list && count, "pm_deposit_plan_brick_list: null pointer");
  if (plan->mode == PM_DEPOSIT_BRICKED) {
    pm::depbrick::node_list(plan->brick, list, count, max_count);
    return PM_OK;
  }
  *list = plan->bricks_valid ? plan->brick_list : nullptr;
  *count = plan->bricks_valid ? plan->brick_count : nullptr;
  if (max_count) *max_count = nbricks_of(plan->n);
  return PM_OK;
}

pm_status pm_deposit_plan_order(pm_deposit_plan* plan, const int32_t** order) {
  PM_REQUIRE(plan && order, "pm_deposit_plan_order: null pointer");
  *order = plan->mode == PM_DEPOSIT_BRICKED ? pm::depbrick::order(plan->brick) : plan->order;
  return PM_OK;
}

pm_status pm_deposit_plan_mark_ordered(pm_deposit_plan* plan) {
  pm::clear_error();
  PM_REQUIRE(plan, "pm_deposit_plan_mark_ordered: null plan");
  PM_REQUIRE(plan->mode == PM_DEPOSIT_BRICKED,
             "pm_deposit_plan_mark_ordered: needs PM_DEPOSIT_BRICKED mode");
  return pm::depbrick::mark_ordered(plan->brick);
}

pm_status pm_deposit_plan_node_bricks(pm_deposit_plan* plan, const uint8_t** bricks,
                                      int32_t* brick_edge) {
  PM_REQUIRE(plan && bricks, "pm_deposit_plan_node_bricks: null pointer");
  *bricks = plan->bricks_valid ? plan->bricks : nullptr;
  if (brick_edge) *brick_edge = kBrick;
  return PM_OK;
}

pm_status pm_deposit_plan_set_timing(pm_deposit_plan* plan, int32_t enable) {
  PM_REQUIRE(plan, "pm_deposit_plan_set_timing: null plan");
  plan->timing = enable != 0;
  return PM_OK;
}

pm_status pm_deposit_last_timing(pm_deposit_plan* plan, float out_ms[4]) {
  PM_REQUIRE(plan && out_ms, "pm_deposit_last_timing: null pointer");
  if (plan->timing) {
    PM_TRY_HIP(hipEventSynchronize(plan->ev[4]));
    for (int q = 0; q < 4; ++q)
      PM_TRY_HIP(hipEventElapsedTime(&plan->last_ms[q], plan->ev[q], plan->ev[q + 1]));
  }
  for (int q = 0; q < 4; ++q) out_ms[q] = plan->last_ms[q];
  return PM_OK;
}

}  // extern "C"

namespace {

// The deposit; keys_in (optional) holds the particles' keys already, so the key pass is
// skipped and the first sort pass reads them where they lie.
pm::depbrick::SortBuffers sort_buffers(pm_deposit_plan* p) {
  pm::depbrick::SortBuffers sb;
  sb.keys[0] = p->keys[0];
  sb.keys[1] = p->keys[1];
  sb.vals[0] = p->vals[0];
  sb.vals[1] = p->vals[1];
  sb.hist = p->hist;
  sb.digit = p->digit;
  return sb;
}

// bin_always: a PM_DEPOSIT_BRICKED deposit bins the particles again (pm_deposit_cic) rather
// than reusing the last binning (pm_deposit_cic_step without PM_STEP_BIN).
pm_status deposit_cic(pm_deposit_plan* plan, const float* px, const float* py, const float* pz,
                      const float* mass, int64_t np, float dx, float* rho,
                      const uint32_t* keys_in, uint32_t* keys_out, int32_t* perm_out,
                      hipStream_t stream, bool bin_always = true, bool rho_kept = false) {
  PM_REQUIRE(plan, "pm_deposit_cic: null plan");
  PM_REQUIRE(rho, "pm_deposit_cic: null rho");
  PM_REQUIRE(np >= 0, "pm_deposit_cic: negative particle count");
  PM_REQUIRE(dx > 0.0f, "pm_deposit_cic: dx must be positive");
  if (np > plan->cap) {
    pm::set_error("pm_deposit_cic: " + std::to_string(np) + " particles exceed the plan capacity " +
                  std::to_string(plan->cap));
    return PM_ERR_CAPACITY;
  }
  PM_REQUIRE(np == 0 || (px && py && pz && mass), "pm_deposit_cic: null particle array");
  plan->bricks_valid = plan->mode == PM_DEPOSIT_SORTED;
  plan->order = nullptr;
  if (plan->mode == PM_DEPOSIT_BRICKED) {
    PM_REQUIRE(!keys_in && !keys_out && !perm_out,
               "pm_deposit_cic: PM_DEPOSIT_BRICKED mode takes no keys or permutation");
    return pm::depbrick::deposit(plan->brick, sort_buffers(plan), px, py, pz, mass, np, dx, rho,
                                 bin_always, rho_kept, stream);
  }
  if (plan->mode == PM_DEPOSIT_FIXED) {
    PM_REQUIRE(!perm_out, "pm_deposit_cic: the particle order needs PM_DEPOSIT_SORTED mode");
    return pm::depfix::deposit(px, py, pz, mass, np, plan->n, dx, rho, keys_out, plan->acc64,
                               plan->maxbits, stream);
  }
  const int n = plan->n;
  const int64_t nodes = static_cast<int64_t>(n) * n * n;
  const int64_t m = static_cast<int64_t>(n) + 1;
  const int64_t cells = m * m * m;
  const bool vec2 = n % 2 == 0 && (reinterpret_cast<uintptr_t>(rho) & 7u) == 0;
  const bool timing = plan->timing;
  if (timing) PM_TRY_HIP(hipEventRecord(plan->ev[0], stream));

  {
    PM_TIMED("deposit.fill_map", 4.0 * cells, stream);
    if (plan->map_clean) {
      reset_listed_kernel<<<kPersistentBlocks, pm::kBlock, 0, stream>>>(
          plan->map, n, plan->bricks, plan->brick_list, plan->brick_count, plan->big_count);
    } else {
      reset_map_kernel<<<pm::blocks_for(cells), pm::kBlock, 0, stream>>>(
          plan->map, cells, plan->bricks, nbricks_of(n), plan->big_count);
    }
    PM_CHECK_LAUNCH("reset_map_kernel");
    plan->map_clean = true;  // holds again once this deposit's marks are written
  }
  if (np == 0) {
    if (timing) {
      for (int q = 1; q < 5; ++q) PM_TRY_HIP(hipEventRecord(plan->ev[q], stream));
    }
    {
      PM_TIMED("deposit.assemble", 8.0 * nodes, stream);
      zero_rho_kernel<<<pm::blocks_for((nodes + 3) / 4), pm::kBlock, 0, stream>>>(
          rho, nodes, (reinterpret_cast<uintptr_t>(rho) & 15u) == 0);
      PM_CHECK_LAUNCH("zero_rho_kernel");
      list_bricks_kernel<<<1, kListLanes, 0, stream>>>(plan->bricks, nbricks_of(n),
                                                       plan->brick_list, plan->brick_count);
      PM_CHECK_LAUNCH("list_bricks_kernel");
      assemble_kernel<<<kPersistentBlocks, 256, 0, stream>>>(
          plan->map, plan->sums, n, rho, vec2, plan->brick_list, plan->brick_count);
      PM_CHECK_LAUNCH("assemble_kernel");
    }
    return PM_OK;
  }

  // 1. keys
  const float inv_vol = 1.0f / ((dx * dx) * dx);
  if (!keys_in) {
    PM_TIMED("deposit.key", 16.0 * np, stream);
    key_kernel<<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(
        px, py, pz, np, n, dx, plan->sentinel, plan->keys[0], keys_out);
    PM_CHECK_LAUNCH("key_kernel");
  } else if (keys_out) {
    PM_TRY_HIP(hipMemcpyAsync(keys_out, keys_in, sizeof(uint32_t) * np,
                              hipMemcpyDeviceToDevice, stream));
  }
  if (timing) PM_TRY_HIP(hipEventRecord(plan->ev[1], stream));

  // 2. sort (the first pass sorts particle indices: no index array is read)
  int src = 0;
  {
    pm_status st = pm::radix::sort(keys_in ? keys_in : plan->keys[0], plan->keys, plan->vals,
                                   plan->hist, plan->digit, np, 0, plan->digit_bits,
                                   plan->passes, &src, stream);
    if (st != PM_OK) return st;
  }
  const uint32_t* skey = plan->keys[src];
  const int32_t* sidx = plan->vals[src];
  plan->order = sidx;
  if (perm_out)
    PM_TRY_HIP(hipMemcpyAsync(perm_out, sidx, sizeof(int32_t) * np, hipMemcpyDeviceToDevice,
                              stream));
  if (timing) PM_TRY_HIP(hipEventRecord(plan->ev[2], stream));

  // 3. segmented reduction
  const int64_t chunks = (np + kChunk - 1) / kChunk;
  {
    PM_TIMED("deposit.piece", 24.0 * np, stream);
    const unsigned piece_blocks = static_cast<unsigned>((np + kPiecePos - 1) / kPiecePos);
    piece_kernel<<<piece_blocks, 256, 0, stream>>>(
        skey, sidx, px, py, pz, mass, dx, inv_vol, np, plan->sentinel, chunks, plan->sums,
        plan->map, plan->head, plan->tail, plan->tail_slot, plan->head_end, n, plan->bricks,
        pm::make_fast_div(static_cast<uint32_t>(n) + 1u));
    PM_CHECK_LAUNCH("piece_kernel");
  }
  {
    PM_TIMED("deposit.join", 8.0 * chunks, stream);
    join_kernel<<<pm::blocks_for(chunks), pm::kBlock, 0, stream>>>(
        skey, np, chunks, plan->tail_slot, plan->head_end, plan->head, plan->tail, plan->sums,
        plan->cap, plan->map, plan->big, plan->big_count, plan->mid_pieces);
    PM_CHECK_LAUNCH("join_kernel");
  }
  {
    PM_TIMED("deposit.mid_cell", 0.0, stream);
    const unsigned mid_blocks =
        static_cast<unsigned>(std::min<int64_t>((chunks + 63) / 64, int64_t(4096)));
    mid_cell_kernel<<<mid_blocks, 256, 0, stream>>>(skey, plan->tail_slot, plan->mid_pieces, chunks,
                                              plan->head, plan->tail, plan->sums, plan->map);
    PM_CHECK_LAUNCH("mid_cell_kernel");
  }
  {
    PM_TIMED("deposit.big_cell", 0.0, stream);
    big_scan_kernel<<<1, 256, 0, stream>>>(plan->big, plan->big_count, plan->big_off);
    PM_CHECK_LAUNCH("big_scan_kernel");
    big_batch_kernel<<<512, 256, 0, stream>>>(plan->big, plan->big_count, plan->big_off,
                                              plan->head, plan->tail);
    PM_CHECK_LAUNCH("big_batch_kernel");
    big_merge_kernel<<<64, 256, 0, stream>>>(skey, plan->big, plan->big_count, plan->head,
                                             plan->tail, plan->sums, plan->map);
    PM_CHECK_LAUNCH("big_merge_kernel");
  }
  if (timing) PM_TRY_HIP(hipEventRecord(plan->ev[3], stream));

  // 4. nodes
  {
    PM_TIMED("deposit.assemble", 8.0 * nodes, stream);
    zero_rho_kernel<<<pm::blocks_for((nodes + 3) / 4), pm::kBlock, 0, stream>>>(
          rho, nodes, (reinterpret_cast<uintptr_t>(rho) & 15u) == 0);
      PM_CHECK_LAUNCH("zero_rho_kernel");
      list_bricks_kernel<<<1, kListLanes, 0, stream>>>(plan->bricks, nbricks_of(n),
                                                       plan->brick_list, plan->brick_count);
      PM_CHECK_LAUNCH("list_bricks_kernel");
      assemble_kernel<<<kPersistentBlocks, 256, 0, stream>>>(
          plan->map, plan->sums, n, rho, vec2, plan->brick_list, plan->brick_count);
    PM_CHECK_LAUNCH("assemble_kernel");
  }
  if (timing) PM_TRY_HIP(hipEventRecord(plan->ev[4], stream));
  return PM_OK;
}

}  // namespace

extern "C" {

pm_status pm_deposit_cic(pm_deposit_plan* plan, const float* px, const float* py,
                         const float* pz, const float* mass, int64_t np, float dx, float* rho,
                         uint32_t* keys_out, int32_t* perm_out, hipStream_t stream) {
  return deposit_cic(plan, px, py, pz, mass, np, dx, rho, nullptr, keys_out, perm_out, stream);
}

pm_status pm_deposit_cic_keyed(pm_deposit_plan* plan, const float* px, const float* py,
                               const float* pz, const float* mass, int64_t np, float dx,
                               float* rho, const uint32_t* keys, int32_t* perm_out,
                               hipStream_t stream) {
  PM_REQUIRE(np == 0 || keys, "pm_deposit_cic_keyed: null keys");
  return deposit_cic(plan, px, py, pz, mass, np, dx, rho, keys, nullptr, perm_out, stream);
}

pm_status pm_deposit_cic_step(pm_deposit_plan* plan, const float* px, const float* py,
                              const float* pz, const float* mass, int64_t np, float dx,
                              float* rho, int32_t flags, hipStream_t stream) {
  PM_REQUIRE(plan, "pm_deposit_cic_step: null plan");
  PM_REQUIRE(plan->mode == PM_DEPOSIT_BRICKED, "pm_deposit_cic_step: needs PM_DEPOSIT_BRICKED mode");
  PM_REQUIRE((flags & ~(PM_STEP_BIN | PM_STEP_RHO_KEPT)) == 0, "pm_deposit_cic_step: unknown flags");
  return deposit_cic(plan, px, py, pz, mass, np, dx, rho, nullptr, nullptr, nullptr, stream,
                     (flags & PM_STEP_BIN) != 0, (flags & PM_STEP_RHO_KEPT) != 0);
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// The deposit without a plan: its scratch is a static device array, carved per call.
// ------------------------------------------------------------------------------------

// Every device the library runs on holds one copy of this array until the process ends.
// Host code never passes the array's name to a kernel: it stands for the device array only
// in hipGetSymbolAddress, which gives each device's own address (noalloc_base).
static __device__ __align__(256) unsigned char scratch[PM_DEPOSIT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// The scratch array's address on the current device, looked up once per device.
pm_status noalloc_base(unsigned char** out) {
  int dev = 0;
  PM_TRY_HIP(hipGetDevice(&dev));
  PM_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices,
             "pm_deposit_cic_noalloc: device ordinal beyond the supported 64");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return pm::hip_fail("hipGetSymbolAddress (deposit scratch)", e);
    PM_REQUIRE(ptr, "hipGetSymbolAddress gave no address for the deposit scratch");
    g_noalloc_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_noalloc_ptr[dev];
  return PM_OK;
}

// Bytes the plan's buffers take when each starts on a 256-byte boundary.
int64_t carved_bytes(pm_deposit_plan* p) {
  int64_t off = 0;
  (void)for_each_buffer(p, [&](auto*& ptr, int64_t count, const char*) {
    using T = std::remove_reference_t<decltype(*ptr)>;
    off = (off + 255) & ~int64_t(255);
    off += static_cast<int64_t>(sizeof(T)) * (count > 0 ? count : 1);
    return PM_OK;
  });
  return off;
}

}  // namespace

extern "C" {

pm_status pm_deposit_cic_noalloc(int32_t n, const float* px, const float* py, const float* pz,
                                 const float* mass, int64_t np, float dx, float* rho,
                                 uint32_t* keys_out, int32_t* perm_out, hipStream_t stream) {
  pm::clear_error();
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_deposit_cic_noalloc: grid_size must be in [1, 1024]");
  PM_REQUIRE(np >= 0 && np < (int64_t(1) << 31),
             "pm_deposit_cic_noalloc: particle count must be in [0, 2^31)");
  pm_deposit_plan plan;  // host bookkeeping only; every buffer lies in the scratch array
  init_plan_shape(&plan, n, np > 0 ? np : 1);
  const int64_t need = carved_bytes(&plan);
  if (need > PM_DEPOSIT_NOALLOC_SCRATCH_BYTES) {
    pm::set_error("pm_deposit_cic_noalloc: " + std::to_string(np) + " particles on a " +
                  std::to_string(n) + "^3 grid need " + std::to_string(need) +
                  " bytes of scratch; the static scratch holds " +
                  std::to_string(int64_t(PM_DEPOSIT_NOALLOC_SCRATCH_BYTES)));
    return PM_ERR_CAPACITY;
  }
  unsigned char* base = nullptr;
  pm_status st = noalloc_base(&base);
  if (st != PM_OK) return st;
  int64_t off = 0;
  (void)for_each_buffer(&plan, [&](auto*& ptr, int64_t count, const char*) {
    using T = std::remove_reference_t<decltype(*ptr)>;
    off = (off + 255) & ~int64_t(255);
    ptr = reinterpret_cast<T*>(base + off);
    off += static_cast<int64_t>(sizeof(T)) * (count > 0 ? count : 1);
    return PM_OK;
  });
  // map_clean is false: the deposit resets the whole cell map, the brick marks and the big
  // cell count itself, whatever an earlier call left in the array; no timer records
  pm::TimingPause no_timers;
  return deposit_cic(&plan, px, py, pz, mass, np, dx, rho, nullptr, keys_out, perm_out,
                     stream);
}

}  // extern "C"
