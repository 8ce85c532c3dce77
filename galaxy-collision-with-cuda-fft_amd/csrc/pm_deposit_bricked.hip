// Bricked fixed-point CIC deposit: particles binned by 8^3-cell bricks, each brick's
// 64-bit fixed-point corner sums accumulated in LDS. Replaces reference deposit.py:47-82
// (`_deposit_cic`) like the other modes.
//
// Arithmetic (the contract of PM_DEPOSIT_FIXED, restated by the oracle's
// oracle_deposit_fixed): per particle g = x/dx, base cell floor(g), offsets f, corner term
// t = ((m'*wx)*wy)*wz in float32 with m' = mass*(1/((dx*dx)*dx)); corners outside [0, n)
// and non-finite or non-positive terms dropped; q = round-half-even(t*2^s) (pm_fixed.h);
// node sums of q are exact integers, so the order in which particles arrive does not matter;
// rho = float(double(sum)*2^-s).
//
// Layout. Extended cell e = i+1 in [0, n] per axis (i = base node index); cell brick
// b = e >> 3 (nbc = ceil((n+1)/8) per axis). A cell brick touches nodes [8b-1, 8b+7]; its
// region adds one cell of slack on each side: nodes [8b-2, 8b+8], 11 per axis, 1331
// 64-bit sums.
//
// Binning (when asked for, or when there is none): key = (cell brick, cell in the brick)
// (or a sentinel past the last brick for particles with no corner in the box), a stable
// radix sort of the keys (so a brick's particles are in cell order), brick bounds in the
// sorted order, and a unit table: each brick's positions cut into units of at most kUnit,
// then the positions of the sentinel particles as units of no brick.
//
// Every deposit (three launches):
//   unit pass: workgroups claim units from a queue, the larger first; per unit a workgroup
//     zeroes the 1331 LDS sums, reads the unit's particles (through the binning's order,
//     or in place once the caller has moved the particles into that order), loaded
//     coalesced and passed through LDS so that each lane takes runs of consecutive
//     (cell-ordered) positions, keeps the corner q of up to two cells in registers, adds
//     them to the LDS sums at the end of a round (those many lanes share summed over the
//     wave first), and stores the region's sums for its unit. A particle outside the region
//     (it moved more than one cell out of its brick since the binning, or has no brick)
//     adds its q to a zeroed n^3 64-bit array with global atomics and flags the node bricks
//     it touched: exact, only slower, so the result never depends on how fresh the binning
//     is;
//   brick reduce: the region sums of a brick with several units are added into its first;
//   combine pass: every 8^3 node brick sums, per node, the unit sums of the cell bricks whose
//     region holds the node (the brick itself; the one below for the first layer; the one
//     above for the last two), adds and clears the flagged 64-bit array entries, converts,
//     and stores rho; node bricks with no unit nearby store zeros (or keep the zeros they
//     hold when rho is the last deposit's array).
#include <algorithm>
#include <climits>

#include "pm_cic.h"
#include "pm_common.h"
#include "pm_deposit_bricked.h"
#include "pm_deposit_fixed.h"
#include "pm_fixed.h"
#include "pm_radix.h"
#include "pm_timing.h"

namespace {

constexpr int kCB = 8;                   // extended cells per brick edge
constexpr int kCellBits = 9;             // bits of a cell's index within its brick
constexpr int kR = kCB + 1 + 2;          // region nodes per axis (one cell of slack a side)
constexpr int kRN = kR * kR * kR;        // 1331 sums per region
constexpr int kRegion = kRN + 1;         // stride of the stored regions (16-byte aligned)
constexpr int kUnit = 8192;              // binned positions per unit (at most)
// With N >= 2^(kF64Bits-1) particles a unit's node sums stay below kUnit * 2^(62-b) <= 2^53
// (b = bit length of N), so float64 LDS sums of the integer terms are exact.
constexpr int kF64Bits = 13 + 62 - 53;
static_assert(kUnit == 1 << 13, "kF64Bits assumes 2^13 positions per unit");
constexpr unsigned kPersistent = 2048;   // workgroups of the persistent passes
constexpr int kUnitGroups = 768;         // workgroups of the unit pass (3 per CU stay resident)
static_assert(kCB == 8 && pm::kNodeBrick == 8, "bricks of 8 cells / 8 nodes per axis");

struct Unit {
  int32_t brick;  // cell brick, or -1 for particles with no brick at binning
  int32_t begin;  // first binned position
  int32_t count;  // positions
  int32_t units;  // first unit of a brick: the brick's number of units; others 0 (in the
                  // claim order: the unit's index)
};

__device__ __forceinline__ void decompose(uint32_t b, uint32_t nb, int& x, int& y, int& z) {
  z = static_cast<int>(b % nb);
  y = static_cast<int>((b / nb) % nb);
  x = static_cast<int>(b / (nb * nb));
}

// Layout of a unit's 11^3 region sums in memory: per axis the region coordinate r falls in
// class 0 (r < 2: nodes 8b-2, 8b-1, read by the node brick below), 1 (2 <= r < 10: the
// brick's own 8 node layers) or 2 (r = 10: node 8b+8, read by the node brick above); the
// 27 class blocks are stored one after another, each contiguous, so the node brick that
// reads a block reads whole cache lines of it.
__host__ __device__ constexpr int cls_lo(int c) { return c == 0 ? 0 : (c == 1 ? 2 : 10); }
__host__ __device__ constexpr int cls_n(int c) { return c == 0 ? 2 : (c == 1 ? 8 : 1); }
__host__ __device__ constexpr int cls_of(int r) { return r < 2 ? 0 : (r < 10 ? 1 : 2); }
__host__ __device__ constexpr int block_off(int cx, int cy, int cz) {
  return kR * kR * cls_lo(cx) + kR * cls_n(cx) * cls_lo(cy) + cls_n(cx) * cls_n(cy) * cls_lo(cz);
}
__device__ __forceinline__ int region_slot(int rx, int ry, int rz) {
  const int cx = cls_of(rx), cy = cls_of(ry), cz = cls_of(rz);
  return block_off(cx, cy, cz) +
         ((rx - cls_lo(cx)) * cls_n(cy) + (ry - cls_lo(cy))) * cls_n(cz) + (rz - cls_lo(cz));
}
static_assert(block_off(2, 2, 2) + 1 == kRN, "class blocks tile the region");

// ---------------------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------------------
// Binning key of every particle, (cell brick << kCellBits) | cell in the brick (grid-stride),
// and this workgroup's largest finite positive
// m' = mass*inv_vol as float bits in maxparts[blockIdx.x] (positive floats order like their
// bits), folded by unit_build_kernel.
__global__ void __launch_bounds__(256) brick_key_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ mass, int64_t np, int n, float dx, float inv_vol, uint32_t nbc,
    uint32_t sentinel, uint32_t* __restrict__ keys, unsigned int* __restrict__ maxparts) {
  __shared__ unsigned int s_max[4];
  unsigned int mx = 0u;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < np;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float x = px[p], y = py[p], z = pz[p], m = mass[p];
    int ix = 0, iy = 0, iz = 0;
    float fx, fy, fz;
    const bool in = pm::cic_axis(x, dx, n, ix, fx) && pm::cic_axis(y, dx, n, iy, fy) &&
                    pm::cic_axis(z, dx, n, iz, fz);
    const uint32_t bx = static_cast<uint32_t>(ix + 1) >> 3, by = static_cast<uint32_t>(iy + 1) >> 3,
                   bz = static_cast<uint32_t>(iz + 1) >> 3;
    const uint32_t cell = ((static_cast<uint32_t>(ix + 1) & 7u) << 6) |
                          ((static_cast<uint32_t>(iy + 1) & 7u) << 3) |
                          (static_cast<uint32_t>(iz + 1) & 7u);
    keys[p] = in ? (((bx * nbc + by) * nbc + bz) << kCellBits) | cell : sentinel;
    const float v = m * inv_vol;
    const unsigned int bits = (v > 0.0f && v < __builtin_huge_valf()) ? __float_as_uint(v) : 0u;
    mx = bits > mx ? bits : mx;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned int o = __shfl_xor(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int m4 = s_max[0];
    for (int w = 1; w < 4; ++w) m4 = s_max[w] > m4 ? s_max[w] : m4;
    maxparts[blockIdx.x] = m4;
  }
}

// first[b] / end[b] = the bounds of brick b in the sorted keys (both 0 for absent bricks:
// unit_build_kernel clears what it reads).
__global__ void __launch_bounds__(256) brick_bounds_kernel(const uint32_t* __restrict__ skeys,
                                                           int64_t np, uint32_t ncb,
                                                           int32_t* __restrict__ first,
                                                           int32_t* __restrict__ end) {
  const int64_t j = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (j >= np) return;
  const uint32_t b = skeys[j] >> kCellBits;
  const uint32_t prev = j > 0 ? skeys[j - 1] >> kCellBits : 0xFFFFFFFFu;
  const uint32_t next = j + 1 < np ? skeys[j + 1] >> kCellBits : 0xFFFFFFFFu;
  if (b >= ncb) return;
  if (prev != b) first[b] = static_cast<int32_t>(j);
  if (next != b) end[b] = static_cast<int32_t>(j + 1);
}

// Exclusive scan of one value per thread over a 1024-thread workgroup; total = the sum.
__device__ __forceinline__ uint32_t scan_1024(uint32_t v, uint32_t& total, uint32_t* s_wave) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
  for (int w = 0; w < 16; ++w) {
    const uint32_t x = s_wave[w];
    if (w < wave) before += x;
    total += x;
  }
  __syncthreads();
  return before + incl - v;
}

// Entries per thread and round of the one-workgroup table passes: thread t takes entries
// round + 1024*q + t (q < kSpan), so every load and store of a wave is coalesced; all of a
// thread's loads are issued together (the tables are padded to whole rounds).
constexpr int kSpan = 32;
constexpr int64_t kRound = int64_t(1024) * kSpan;
inline int64_t padded(int64_t count) { return (count + kRound - 1) / kRound * kRound; }

// One workgroup: brick b gets ceil(count/unit) units from unit_first[b]; the sentinel
// positions [binned, np) follow as units of no brick. Clears first/end behind itself.
__global__ void __launch_bounds__(1024) unit_build_kernel(int32_t* __restrict__ first,
                                                          int32_t* __restrict__ end,
                                                          uint32_t ncb, int64_t np, int unit,
                                                          Unit* __restrict__ units,
                                                          int32_t* __restrict__ unit_first,
                                                          int32_t* __restrict__ unit_num,
                                                          uint32_t* __restrict__ unit_count,
                                                          unsigned int* __restrict__ maxbits,
                                                          int maxparts,
                                                          int32_t* __restrict__ multi,
                                                          uint32_t* __restrict__ multi_count) {
  __shared__ uint32_t s_wave[16];
  {
    // maxbits[0] = the largest of the key pass's per-workgroup maxima maxbits[1..]
    unsigned int mx = 0u;
    for (int q = threadIdx.x; q < maxparts; q += blockDim.x)
      mx = maxbits[1 + q] > mx ? maxbits[1 + q] : mx;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned int o = __shfl_xor(mx, off, 64);
      mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned int m = 0u;
      for (int w = 0; w < 16; ++w) m = s_wave[w] > m ? s_wave[w] : m;
      maxbits[0] = m;
    }
    __syncthreads();
  }
  uint32_t units_before = 0, binned = 0, multi_before = 0;  // totals of the earlier rounds
  for (uint32_t round = 0; round < ncb; round += static_cast<uint32_t>(kRound)) {
    const uint32_t b0 = round + threadIdx.x;
    uint32_t nu = 0, nc = 0, nm = 0;
#pragma unroll 8
    for (int q = 0; q < kSpan; ++q) {
      const uint32_t cnt = static_cast<uint32_t>(end[b0 + 1024u * q] - first[b0 + 1024u * q]);
      const uint32_t k = (cnt + unit - 1) / unit;
      nu += k;
      nc += cnt;
      nm += k >= 2 ? 1u : 0u;
    }
    // a thread's bricks take consecutive units (any order of the bricks serves)
    uint32_t round_units, round_binned, round_multi;
    uint32_t base = units_before + scan_1024(nu, round_units, s_wave);
    (void)scan_1024(nc, round_binned, s_wave);
    uint32_t mbase = multi_before + scan_1024(nm, round_multi, s_wave);
    // the next brick's bounds are loaded while this one's units are written
    int32_t fn = first[b0], en = end[b0];
#pragma unroll 1
    for (int q = 0; q < kSpan; ++q) {
      const uint32_t b = b0 + 1024u * q;
      if (b >= ncb) break;
      const int32_t f = fn;
      const uint32_t cnt = static_cast<uint32_t>(en - fn);
      if (q + 1 < kSpan) {
        fn = first[b + 1024u];
        en = end[b + 1024u];
      }
      const uint32_t k = (cnt + unit - 1) / unit;
      unit_first[b] = static_cast<int32_t>(base);
      unit_num[b] = static_cast<int32_t>(k);
      for (uint32_t w = 0; w < k; ++w) {
        Unit un;
        un.brick = static_cast<int32_t>(b);
        un.begin = f + static_cast<int32_t>(w * unit);
        un.count = static_cast<int32_t>(min(static_cast<uint32_t>(unit), cnt - w * unit));
        un.units = w == 0 ? static_cast<int32_t>(k) : 0;
        units[base + w] = un;
      }
      if (k >= 2) multi[mbase++] = static_cast<int32_t>(base);
      base += k;
      if (cnt) {
        first[b] = 0;
        end[b] = 0;
      }
    }
    units_before += round_units;
    binned += round_binned;
    multi_before += round_multi;
  }
  // the particles with no brick
  const int64_t rest = np - static_cast<int64_t>(binned);
  const uint32_t tail = static_cast<uint32_t>((rest + unit - 1) / unit);
  for (uint32_t q = threadIdx.x; q < tail; q += blockDim.x) {
    Unit un;
    un.brick = -1;
    un.begin = static_cast<int32_t>(binned + q * unit);
    un.count = static_cast<int32_t>(min(static_cast<int64_t>(unit), rest - int64_t(q) * unit));
    un.units = 0;
    units[units_before + q] = un;
  }
  if (threadIdx.x == 0) {
    unit_count[0] = units_before + tail;
    multi_count[0] = multi_before;
  }
}

// Bit (dx+1)*9 + (dy+1)*3 + (dz+1) of node_src[B] is set when cell brick B + (dx, dy, dz)
// has units: the sources of node brick B's sums.
__global__ void __launch_bounds__(256) node_src_kernel(const int32_t* __restrict__ unit_num,
                                                       int nbc, int nbn,
                                                       uint32_t* __restrict__ node_src) {
  const int64_t B = pm::xcd_block() * blockDim.x + threadIdx.x;
  const int64_t nnb = static_cast<int64_t>(nbn) * nbn * nbn;
  if (B >= nnb) return;
  int bi, bj, bk;
  decompose(static_cast<uint32_t>(B), static_cast<uint32_t>(nbn), bi, bj, bk);
  uint32_t mask = 0;
  int nums[27];
#pragma unroll
  for (int d = 0; d < 27; ++d) {
    const int x = bi + d / 9 - 1, y = bj + (d / 3) % 3 - 1, z = bk + d % 3 - 1;
    const bool in = x >= 0 && x < nbc && y >= 0 && y < nbc && z >= 0 && z < nbc;
    nums[d] = in ? unit_num[(static_cast<int64_t>(x) * nbc + y) * nbc + z] : 0;
  }
#pragma unroll
  for (int d = 0; d < 27; ++d) mask |= nums[d] > 0 ? (1u << d) : 0u;
  node_src[B] = mask;
}

// One workgroup: list = node bricks with sources (count[1] of them, any order), then the
// others from the back; count[0] = count[1] (the gradient's count, raised to all node
// bricks by a deposit that had particles outside their regions).
__global__ void __launch_bounds__(1024) node_list_kernel(const uint32_t* __restrict__ node_src,
                                                         int64_t nnb, int32_t* __restrict__ list,
                                                         uint32_t* __restrict__ count) {
  __shared__ uint32_t s_act, s_idle;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) s_act = s_idle = 0u;
  __syncthreads();
  for (int64_t round = 0; round < nnb; round += kRound) {
    const int64_t lo = round + threadIdx.x;
    uint32_t m[kSpan];
#pragma unroll
    for (int q = 0; q < kSpan; ++q) m[q] = node_src[lo + 1024 * q];
    uint32_t bits = 0, act = 0, idle = 0;
#pragma unroll
    for (int q = 0; q < kSpan; ++q) {
      const bool valid = lo + 1024 * q < nnb;
      bits |= (valid && m[q]) ? (1u << q) : 0u;
      act += (valid && m[q]) ? 1u : 0u;
      idle += (valid && !m[q]) ? 1u : 0u;
    }
    uint32_t ia = act, ii = idle;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t ta = __shfl_up(ia, off, 64), ti = __shfl_up(ii, off, 64);
      if (lane >= off) {
        ia += ta;
        ii += ti;
      }
    }
    uint32_t wa = 0, wi = 0;
    if (lane == 63) {
      wa = atomicAdd(&s_act, ia);
      wi = atomicAdd(&s_idle, ii);
    }
    uint32_t pa = __shfl(wa, 63, 64) + ia - act, pi = __shfl(wi, 63, 64) + ii - idle;
#pragma unroll
    for (int q = 0; q < kSpan; ++q) {
      const int64_t B = lo + 1024 * q;
      if (B >= nnb) continue;
      if ((bits >> q) & 1u) {
        list[pa++] = static_cast<int32_t>(B);
      } else {
        list[nnb - 1 - pi++] = static_cast<int32_t>(B);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    count[0] = s_act;
    count[1] = s_act;
  }
}

// Source descriptors of the listed node bricks with sources, read by the combine pass in
// one step: desc[e*27 + d] = (first unit, units) of cell brick B + offset d of node brick
// B = list[e], e < count[1] ((0, 0) where d has no units).
__global__ void __launch_bounds__(256) node_desc_kernel(const int32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ count,
                                                        const uint32_t* __restrict__ node_src,
                                                        const int32_t* __restrict__ unit_first,
                                                        const int32_t* __restrict__ unit_num,
                                                        int nbc, int nbn,
                                                        int2* __restrict__ desc) {
  const int64_t t = pm::xcd_block() * blockDim.x + threadIdx.x;
  const int64_t e = t / 27;
  const int d = static_cast<int>(t - e * 27);
  if (e >= count[1]) return;
  const uint32_t B = static_cast<uint32_t>(list[e]);
  int2 out = make_int2(0, 0);
  if ((node_src[B] >> d) & 1u) {
    int bi, bj, bk;
    decompose(B, static_cast<uint32_t>(nbn), bi, bj, bk);
    const int64_t src = (static_cast<int64_t>(bi + d / 9 - 1) * nbc + (bj + (d / 3) % 3 - 1)) * nbc +
                        (bk + d % 3 - 1);
    // the reduce pass leaves a brick's summed region in its first unit
    out = make_int2(unit_first[src], unit_num[src] > 0 ? 1 : 0);
  }
  desc[t] = out;
}

// ---------------------------------------------------------------------------------------
// unit pass
// ---------------------------------------------------------------------------------------
// The LDS sums: float64 when every unit's sums stay exact integers below 2^53 (at most
// kUnit terms of at most 2^(62-b) each: b >= 22), else 64-bit integers. Either way every
// addition is exact, so the order of the adds does not matter.
template <bool F64>
struct Sum;
template <>
struct Sum<true> {
  using T = double;
  // T: a term already scaled by 2^s (exactly, see unit_deposit_kernel): round it there
  static __device__ __forceinline__ T folded(float t) { return static_cast<double>(rintf(t)); }
  // t * 2^s is exact in float32 for s > 0 (a power-of-two scaling that stays below 2^62),
  // so rounding it there gives llrint(ldexp((double)t, s))
  static __device__ __forceinline__ T scaled(float t, int s) {
    return static_cast<double>(rintf(__builtin_amdgcn_ldexpf(t, s)));
  }
  static __device__ __forceinline__ unsigned long long u64(T v) {
    return static_cast<unsigned long long>(v);
  }
};
template <>
struct Sum<false> {
  using T = unsigned long long;
  static __device__ __forceinline__ T folded(float t) {
    return static_cast<unsigned long long>(rintf(t));
  }
  static __device__ __forceinline__ T scaled(float t, int s) { return pm::fixed_q(t, s); }
  static __device__ __forceinline__ unsigned long long u64(T v) { return v; }
};

// Lane swaps inside a row of 16 lanes (DPP: VALU moves, no LDS round trip): partner
// lane^1 and lane^2 (quad permutations), 7-i within each 8 lanes (half-row mirror) and
// i+8 mod 16 (row rotation by 8).
constexpr int kDppXor1 = 0xB1;    // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;    // quad_perm [2,3,0,1]
constexpr int kDppMirror8 = 0x141;
constexpr int kDppRor8 = 0x128;

template <int kCtrl>
__device__ __forceinline__ uint32_t dpp32(uint32_t v) {
  // all lanes take part and every source lane exists, so no old value is merged in
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(v), kCtrl, 0xF, 0xF, true));
}

template <int kCtrl, typename T>
__device__ __forceinline__ T dpp64(T v) {
  static_assert(sizeof(T) == 8, "64-bit values");
  uint64_t bits;
  __builtin_memcpy(&bits, &v, 8);
  const uint32_t lo = dpp32<kCtrl>(static_cast<uint32_t>(bits));
  const uint32_t hi = dpp32<kCtrl>(static_cast<uint32_t>(bits >> 32));
  bits = (static_cast<uint64_t>(hi) << 32) | lo;
  T out;
  __builtin_memcpy(&out, &bits, 8);
  return out;
}

// v from lane src (all lanes take part).
template <typename T>
__device__ __forceinline__ T shfl64(T v, int src) {
  uint64_t bits;
  __builtin_memcpy(&bits, &v, 8);
  const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(bits), src, 64));
  const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(bits >> 32), src, 64));
  bits = (static_cast<uint64_t>(hi) << 32) | lo;
  T out;
  __builtin_memcpy(&out, &bits, 8);
  return out;
}

// Sums of the 8 corner values over each group of 8 lanes, reduce-scattered: lane 8g + j
// returns group g's sum for corner j (partners 7-j, then j^2, then j^1 inside the group).
// Every lane of the wave must take part.
template <typename T>
__device__ __forceinline__ T group_sums(const T (&v)[8], int lane) {
  const bool hi = lane & 4, mid = lane & 2, lo = lane & 1;
  T w4[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const T send = hi ? v[k] : v[4 + k];
    const T keep = hi ? v[4 + k] : v[k];
    w4[k] = keep + dpp64<kDppMirror8>(send);
  }
  T w2[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const T send = mid ? w4[k] : w4[2 + k];
    const T keep = mid ? w4[2 + k] : w4[k];
    w2[k] = keep + dpp64<kDppXor2>(send);
  }
  return (lo ? w2[1] : w2[0]) + dpp64<kDppXor1>(lo ? w2[0] : w2[1]);
}

// Region offset of corner c of a cell.
__device__ __forceinline__ int corner_off(int c) {
  return ((c >> 2) * kR + ((c >> 1) & 1)) * kR + (c & 1);
}

// Adds the 8 corner sums of the lanes with `flushing` set to their cell `cur`. Lanes take
// the binned positions in cell order side by side, so many lanes flush the same cell at
// once: up to kFlushGroups such groups of kFlushMin or more lanes (the cells of the lowest
// and the highest flushing lane first) are summed over the wave (one add per corner), the
// rest add per lane. Every lane of the wave must call this.
constexpr int kFlushGroups = 6;
constexpr int kFlushMin = 8;
template <typename T>
__device__ __forceinline__ void flush_sums(T* __restrict__ acc, int cur, const T (&sum)[8],
                                           bool flushing, int lane) {
  uint64_t fl = __ballot(flushing);
#pragma unroll 1
  for (int it = 0; it < kFlushGroups && __popcll(fl) >= kFlushMin; ++it) {
    const int ka = __builtin_amdgcn_readlane(cur, __builtin_ctzll(fl));
    const int kz = __builtin_amdgcn_readlane(cur, 63 - __builtin_clzll(fl));
    const uint64_t ma = __ballot(cur == ka) & fl, mz = __ballot(cur == kz) & fl;
    const int k0 = __popcll(mz) > __popcll(ma) ? kz : ka;
    const uint64_t same = __popcll(mz) > __popcll(ma) ? mz : ma;
    if (__popcll(same) < kFlushMin) break;
    const bool me = (same >> lane) & 1ull;
    T w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = me ? sum[c] : T(0);
    T z = group_sums(w, lane);
    z = z + dpp64<kDppRor8>(z);
    // rows 0..3 hold their sums in lanes 0..7: add the rows pairwise across the wave
    z = z + shfl64(z, lane ^ 16);
    z = z + shfl64(z, lane ^ 32);
    if (lane < 8) atomicAdd(&acc[k0 + corner_off(lane)], z);
    fl &= ~same;
  }
  if ((fl >> lane) & 1ull) {
#pragma unroll
    for (int c = 0; c < 8; ++c) atomicAdd(&acc[cur + corner_off(c)], sum[c]);
  }
}

// A particle outside its unit's region: exact global adds of its 8 terms, and the node
// bricks it touched flagged for the combine pass. Rare, so kept out of the unit pass's body.
__device__ __noinline__ void add_outside(const unsigned long long (&q)[8], int ix, int iy, int iz,
                                         int n, int nbn, unsigned long long* __restrict__ esc,
                                         uint8_t* __restrict__ eflag,
                                         uint32_t* __restrict__ esc_flag) {
  const int64_t n2 = static_cast<int64_t>(n) * n;
  for (int c = 0; c < 8; ++c) {
    const int jx = ix + (c >> 2), jy = iy + ((c >> 1) & 1), jz = iz + (c & 1);
    if (!q[c] || jx < 0 || jx >= n || jy < 0 || jy >= n || jz < 0 || jz >= n) continue;
    atomicAdd(&esc[jx * n2 + static_cast<int64_t>(jy) * n + jz], q[c]);
    eflag[((jx >> 3) * nbn + (jy >> 3)) * static_cast<int64_t>(nbn) + (jz >> 3)] = 1;
    *esc_flag = 1u;
  }
}

// Buffer resource over `count` 4-byte values at p (binned positions stay below 2^30, so
// byte offsets fit 32 bits): loads through it take one 32-bit offset register and return 0
// past the end instead of reading beyond it. Dword 3 is the gfx9 raw-buffer format.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, int64_t count) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), static_cast<short>(0),
                                           static_cast<int>(static_cast<uint32_t>(count * 4)),
                                           0x00020000);
}
__device__ __forceinline__ uint32_t ldb(__amdgpu_buffer_rsrc_t r, uint32_t index) {
  return static_cast<uint32_t>(__builtin_amdgcn_raw_buffer_load_b32(r, index * 4u, 0, 0));
}
__device__ __forceinline__ float ldbf(__amdgpu_buffer_rsrc_t r, uint32_t index) {
  return __uint_as_float(ldb(r, index));
}

// A wave takes its positions a phase of kPhasePos at a time: loaded coalesced (lane l reads
// positions l, l+64, ... of the phase), stored to the wave's LDS stage, and read back so
// that lane l takes R <= kPhaseRun consecutive positions (R = the phase's positions / 64,
// rounded up), which in a brick's cell order keeps a lane in one cell wherever cells hold
// more than a few particles. A round gives each wave kWavePos positions; the last round of a
// unit (a small unit's only one) splits its remainder evenly over the 4 waves instead, so a
// unit of a few hundred positions takes all waves with runs of one to four positions rather
// than one wave with runs of eight. A lane keeps the 8 corner sums of up to two cells in
// registers over a round of kPhases phases; they go to the LDS sums at the end of the round,
// and a slot taken over by a third cell goes there at once, those that many lanes share
// summed over the wave first (flush_sums).
constexpr int kPhaseRun = 8;                   // consecutive positions per lane and phase, at most
constexpr int kPhasePos = 64 * kPhaseRun;      // positions per wave and phase
constexpr int kPhases = 2;                     // phases per round
constexpr int kWavePos = kPhases * kPhasePos;  // positions per wave and round
constexpr int kRoundPos = 4 * kWavePos;        // positions per workgroup and round
constexpr int kGroup = 4;                      // positions per lane processed together
// Stage layout: position p of a phase at p + p/8 (one pad word per 8), so the 32 lanes of
// each half-wave read their runs (lanes 8l + i, slots 9l + i) from 32 distinct banks; the
// coalesced stores (lanes l + 64j) meet at most two to a bank.
static_assert(kPhaseRun == 8, "the stage padding is laid out for runs of 8");
constexpr int kStageStride = kPhasePos + kPhasePos / 8;  // floats per array and wave
__device__ __forceinline__ int stage_slot(int p) { return p + (p >> 3); }
// Positions per wave in a round with rem positions left: kWavePos in a full round, else the
// remainder split over the 4 waves in multiples of 64 (4 * span >= rem).
__device__ __forceinline__ int wave_span(int rem) {
  return rem >= kRoundPos ? kWavePos : (((rem + 3) >> 2) + 63) & ~63;
}
constexpr int kFoldMax = 64;              // largest scale exponent folded into m'
constexpr int kUnitThreads = 256;         // threads of a unit pass workgroup
typedef uint32_t U32x2 __attribute__((__vector_size__(8)));

// The unit loop of unit_deposit_kernel. FOLD (s <= kFoldMax): the scale 2^s is applied to
// m' once, M = m' * 2^s, and the terms ((M*wx)*wy)*wz are rounded directly. For s <= 64
// this gives the rounded ((m'*wx)*wy)*wz * 2^s of the definition bit for bit: scaling by a
// power of two commutes with every rounding while no value of either product chain is
// below 2^-126, and once one is, both chains' terms stay below 2^(s-126) <= 2^-62 (or below
// 2^-126 for the scaled chain when s < 0) and round to 0.
template <bool F64, bool IDX, bool FOLD>
__device__ __forceinline__ void unit_loop(
    typename Sum<F64>::T* __restrict__ acc, float* __restrict__ stage, int32_t* __restrict__ claim,
    const uint16_t* __restrict__ inv_slot,
    const Unit* __restrict__ work, int32_t nu,
    uint32_t* __restrict__ queue, const int32_t* __restrict__ idx, const float* __restrict__ px,
    const float* __restrict__ py, const float* __restrict__ pz, const float* __restrict__ mass,
    int64_t np, int n, float dx, float inv_vol, int nbc, int nbn, int s,
    unsigned long long* __restrict__ partial, unsigned long long* __restrict__ esc,
    uint8_t* __restrict__ eflag, uint32_t* __restrict__ esc_flag) {
  using S = Sum<F64>;
  using T = typename S::T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const float nf = static_cast<float>(n);
  const float r1 = pm::quot_recip(dx);
  const bool fast = pm::quot_dx_ok(dx);
  const float scale = FOLD ? __builtin_amdgcn_ldexpf(1.0f, s) : 1.0f;  // 2^s, normal for s >= -97
  // a phase's particles of unit U as the coalesced loads see them (positions past the
  // unit's end read its last particle, unused)
  const __amdgpu_buffer_rsrc_t bx = rsrc(px, np), by = rsrc(py, np), bz = rsrc(pz, np),
                               bm = rsrc(mass, np), bi = rsrc(idx, IDX ? np : 0);
  auto load = [&](const Unit& U, int start, float4 (&p)[kPhaseRun]) {
    uint32_t sl[kPhaseRun];
#pragma unroll
    for (int j = 0; j < kPhaseRun; ++j) {
      const uint32_t q = static_cast<uint32_t>(U.begin + min(start + lane + 64 * j, U.count - 1));
      sl[j] = IDX ? ldb(bi, q) : q;
    }
#pragma unroll
    for (int j = 0; j < kPhaseRun; ++j)
      p[j] = make_float4(ldbf(bx, sl[j]), ldbf(by, sl[j]), ldbf(bz, sl[j]), ldbf(bm, sl[j]));
  };
  // the first position of this wave in the round at `base` of a unit of `count` positions
  auto first_pos = [&](int count, int base) { return base + wave * wave_span(count - base); };
  float* const sx = stage + wave * 4 * kStageStride;
  float* const sy = sx + kStageStride;
  float* const sz = sy + kStageStride;
  float* const sm = sz + kStageStride;
  // Units in the work order: workgroup w takes positions w and w + G (G workgroups) of it,
  // then claims positions from 2G on through the queue, at each unit's start the one after
  // the next (the claim's result is only needed at the unit's end; claims spread over the
  // launch instead of all workgroups meeting on one counter at its start). A unit's record
  // (one load: the work order holds the records) is loaded one unit ahead, and a wave's
  // first phase of a unit while it finishes the unit before, so none of these loads'
  // latency adds to a unit's time.
  const int32_t G = static_cast<int32_t>(gridDim.x);
  int32_t cur = static_cast<int32_t>(blockIdx.x), nxt = cur + G;
  Unit un_next = cur < nu ? work[cur] : Unit{};
  float4 ph[kPhaseRun];
  if (cur < nu) load(un_next, first_pos(un_next.count, 0), ph);
  while (cur < nu) {
    // the record is the same in every lane: held in scalar registers for the unit (its
    // load was issued a unit ago)
    Unit un;
    un.brick = __builtin_amdgcn_readfirstlane(un_next.brick);
    un.begin = __builtin_amdgcn_readfirstlane(un_next.begin);
    un.count = __builtin_amdgcn_readfirstlane(un_next.count);
    un.units = __builtin_amdgcn_readfirstlane(un_next.units);
    const int32_t u = un.units;
    const bool more = nxt < nu;
    if (more) un_next = work[nxt];
    uint32_t after = 0u;
    // (the claim's result is first used at the unit's end, so no wait for it here)
    if (tid == 0) after = atomicAdd(queue, 1u);
    bool pref = false;  // ph holds the next unit's first phase
    const bool home = un.brick >= 0;
    int ox = INT_MIN / 2, oy = INT_MIN / 2, oz = INT_MIN / 2;  // region origin (nodes)
    if (home) {
      int bx, by, bz;
      decompose(static_cast<uint32_t>(un.brick), static_cast<uint32_t>(nbc), bx, by, bz);
      ox = kCB * bx - 2;
      oy = kCB * by - 2;
      oz = kCB * bz - 2;
    }
    for (int q = tid; q < kRN; q += blockDim.x) acc[q] = T(0);
    __syncthreads();
    for (int base = 0; base < un.count; base += kRoundPos) {
      // the wave's positions [lo, hi) of this round
      const int span = wave_span(un.count - base);
      const int lo = base + wave * span, hi = min(lo + span, un.count);
      if (lo >= hi) continue;  // nothing of this round for the wave
      // two cells (region indices, -1 when free) with their 8 corner sums: the lane's cell,
      // and the cell past a cell boundary of its runs; a third cell takes slot 1 after
      // adding its sums
      int k0 = -1, k1 = -1;
      T s0[8], s1[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) s0[c] = s1[c] = T(0);
      // (not unrolled: the unrolled body of all kPhases phases outgrew the instruction cache)
#pragma unroll 1
      for (int h = 0; h < kPhases; ++h) {
        const int w0 = lo + h * kPhasePos;  // the phase's first offset
        if (w0 >= hi) break;
        const int m = min(kPhasePos, hi - w0);  // the phase's positions
        const int run = (m + 63) >> 6;          // positions per lane, 1..kPhaseRun
        // the stage is free: the wave's reads of the last phase were issued before these
        // writes, and LDS operations of one wave complete in order
#pragma unroll
        for (int j = 0; j < kPhaseRun; ++j) {
          const int t = stage_slot(64 * j + lane);
          sx[t] = ph[j].x;
          sy[t] = ph[j].y;
          sz[t] = ph[j].z;
          sm[t] = ph[j].w;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the next phase's loads (this round's next phase, the next round's first, or the
        // next unit's first) are issued before this phase's work
        if (w0 + kPhasePos < hi || base + kRoundPos < un.count) {
          load(un, w0 + kPhasePos < hi ? w0 + kPhasePos : first_pos(un.count, base + kRoundPos),
               ph);
        } else if (more && !pref) {
          load(un_next, first_pos(un_next.count, 0), ph);
          pref = true;
        }
        // one binned position of this lane (every lane of the wave calls it together)
        auto body = [&](const float4& pp, bool valid) {
          const float x = pp.x, y = pp.y, z = pp.z;
          float gx = pm::quot(x, dx, r1), gy = pm::quot(y, dx, r1), gz = pm::quot(z, dx, r1);
          const bool slow = !fast || !(fabsf(x) >= 0x1p-60f) || !(fabsf(y) >= 0x1p-60f) ||
                            !(fabsf(z) >= 0x1p-60f);
          if (slow) {
            gx = x / dx;
            gy = y / dx;
            gz = z / dx;
          }
          const bool in = valid && gx >= -1.0f && gx < nf && gy >= -1.0f && gy < nf &&
                          gz >= -1.0f && gz < nf;
          const float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
          const int ix = in ? static_cast<int>(flx) : 0, iy = in ? static_cast<int>(fly) : 0,
                    iz = in ? static_cast<int>(flz) : 0;
          const float fx = gx - flx, fy = gy - fly, fz = gz - flz;
          const float mp = pp.w * inv_vol;
          // m' <= 0 or not finite: every term is dropped (t <= 0 or not finite)
          const bool good = in && mp > 0.0f && mp < __builtin_huge_valf();
          const int rx = ix - ox, ry = iy - oy, rz = iz - oz;
          const bool reg = good && home && static_cast<unsigned>(rx) <= 9u &&
                           static_cast<unsigned>(ry) <= 9u && static_cast<unsigned>(rz) <= 9u;
          const float gx0 = 1.0f - fx, gy0 = 1.0f - fy, gz0 = 1.0f - fz;
          if (good && !reg) {
            // t_c = ((m'*wx)*wy)*wz, the products shared between corners
            const float a0 = mp * gx0, a1 = mp * fx;
            const float bxy[4] = {a0 * gy0, a0 * fy, a1 * gy0, a1 * fy};
            unsigned long long q[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) q[c] = pm::fixed_q(bxy[c >> 1] * ((c & 1) ? fz : gz0), s);
            add_outside(q, ix, iy, iz, n, nbn, esc, eflag, esc_flag);
          }
          // slot of the particle's cell: slot 0 when it holds that cell (or is free), else
          // slot 1, whose cell, when another, first goes to the LDS sums; the lanes doing so
          // at one step mostly share that cell, so it is summed over the wave (every lane
          // takes part in flush_sums)
          const int key = (rx * kR + ry) * kR + rz;
          const bool h0 = reg && (key == k0 || k0 < 0);
          const bool h1 = reg && !h0;
          const bool evict = h1 && k1 >= 0 && key != k1;
          if (__builtin_amdgcn_ballot_w64(evict) != 0ull) {
            flush_sums(acc, k1, s1, evict, lane);
            if (evict) {
#pragma unroll
              for (int c = 0; c < 8; ++c) s1[c] = T(0);
            }
          }
          k0 = h0 ? key : k0;
          k1 = h1 ? key : k1;
          if (reg) {
            const float m2 = FOLD ? mp * scale : mp;
            const float a0 = m2 * gx0, a1 = m2 * fx;
            const float bxy[4] = {a0 * gy0, a0 * fy, a1 * gy0, a1 * fy};
            T v[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
              const float t = bxy[c >> 1] * ((c & 1) ? fz : gz0);
              v[c] = FOLD ? S::folded(t) : S::scaled(t, s);
            }
            if (h0) {
#pragma unroll
              for (int c = 0; c < 8; ++c) s0[c] = s0[c] + v[c];
            }
            if (h1) {
#pragma unroll
              for (int c = 0; c < 8; ++c) s1[c] = s1[c] + v[c];
            }
          }
        };
        if (run == 1) {
          // at most 64 positions: one per lane (a group of kGroup would compute three unused)
          const int t = stage_slot(lane);
          body(make_float4(sx[t], sy[t], sz[t], sm[t]), lane < m);
        } else {
#pragma unroll 1
          for (int g = 0; g < run; g += kGroup) {
            float4 grp[kGroup];
#pragma unroll
            for (int i = 0; i < kGroup; ++i) {
              const int t = stage_slot(run * lane + g + i);  // run * lane + g + i < kPhasePos
              grp[i] = make_float4(sx[t], sy[t], sz[t], sm[t]);
            }
#pragma unroll
            for (int i = 0; i < kGroup; ++i) body(grp[i], g + i < run && run * lane + g + i < m);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      flush_sums(acc, k0, s0, k0 >= 0, lane);
      flush_sums(acc, k1, s1, k1 >= 0, lane);
    }
    if (more && !pref) load(un_next, first_pos(un_next.count, 0), ph);
    if (tid == 0) claim[0] = static_cast<int32_t>(after + 2u * static_cast<uint32_t>(G));
    __syncthreads();
    if (home) {
      // in stored order (consecutive lanes write consecutive words) through a buffer
      // resource over the unit's stored region: 32-bit offsets, every LDS read issued
      // before the first store
      const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc(
          partial + static_cast<int64_t>(u) * kRegion, static_cast<short>(0), kRegion * 8,
          0x00020000);
      constexpr int kOut = (kRN + kUnitThreads - 1) / kUnitThreads;
      int qi[kOut];
#pragma unroll
      for (int k = 0; k < kOut; ++k) {
        const int t = tid + k * kUnitThreads;
        qi[k] = t < kRN ? inv_slot[t] : 0;
      }
      T val[kOut];
#pragma unroll
      for (int k = 0; k < kOut; ++k) val[k] = acc[qi[k]];
#pragma unroll
      for (int k = 0; k < kOut; ++k) {
        const int t = tid + k * kUnitThreads;
        if (t < kRN) {
          const unsigned long long w = S::u64(val[k]);
          U32x2 w2;
          w2[0] = static_cast<uint32_t>(w);
          w2[1] = static_cast<uint32_t>(w >> 32);
          __builtin_amdgcn_raw_buffer_store_b64(w2, rp, t * 8, 0, 0);
        }
      }
    }
    const int32_t claimed = claim[0];
    __syncthreads();  // the next unit zeroes acc and rewrites claim[0]
    cur = nxt;
    nxt = claimed;
  }
}

// IDX: the particles are read through the binning's order idx (else in place). Workgroups
// claim units in the order `work` from queue (zero at the launch) until all nu are taken.
template <bool F64, bool IDX>
__global__ void __launch_bounds__(kUnitThreads, 3) unit_deposit_kernel(
    const Unit* __restrict__ work, const uint32_t* __restrict__ unit_count,
    uint32_t* __restrict__ queue,
    const int32_t* __restrict__ idx, const float* __restrict__ px, const float* __restrict__ py,
    const float* __restrict__ pz, const float* __restrict__ mass, int64_t np, int n, float dx,
    float inv_vol, int nbc, int nbn, const unsigned int* __restrict__ maxbits, int bl,
    unsigned long long* __restrict__ partial, unsigned long long* __restrict__ esc,
    uint8_t* __restrict__ eflag, uint32_t* __restrict__ esc_any, int parity) {
  __shared__ typename Sum<F64>::T acc[kRN];
  __shared__ float stage[4 * 4 * kStageStride];
  __shared__ int32_t claim[1];
  // region index of every stored word (visible after the first unit's barrier)
  __shared__ uint16_t inv_slot[kRN + 1];
  for (int q = threadIdx.x; q < kRN; q += blockDim.x)
    inv_slot[region_slot(q / (kR * kR), (q / kR) % kR, q % kR)] = static_cast<uint16_t>(q);
  const int s = pm::fixed_scale(*maxbits, bl);
  const int32_t nu = static_cast<int32_t>(*unit_count);
  if (s <= kFoldMax)
    unit_loop<F64, IDX, true>(acc, stage, claim, inv_slot, work, nu, queue + parity, idx,
                              px, py, pz, mass, np, n, dx, inv_vol, nbc, nbn, s, partial, esc,
                              eflag, esc_any + parity);
  else
    unit_loop<F64, IDX, false>(acc, stage, claim, inv_slot, work, nu, queue + parity, idx,
                               px, py, pz, mass, np, n, dx, inv_vol, nbc, nbn, s, partial, esc,
                               eflag, esc_any + parity);
}

// One workgroup: the order in which the unit pass claims units, work[0, nu), as the unit
// records with their index in the 4th field: by size class floor(log2(count)), larger
// classes first (a unit takes a fixed few microseconds plus a time proportional to its
// positions), each class in table order, so that the last claims are the shortest. The
// order only balances the workgroups: every unit's sums are its own.
constexpr int kOrderClasses = 14;  // floor(log2(count)) for counts up to kUnit = 2^13
static_assert(kUnit == 1 << (kOrderClasses - 1), "one class per power of two up to kUnit");
__device__ __forceinline__ int order_class(int32_t count) {
  return count > 0 ? 31 - __builtin_clz(static_cast<uint32_t>(count)) : 0;
}
__global__ void __launch_bounds__(1024) unit_order_kernel(const Unit* __restrict__ units,
                                                          const uint32_t* __restrict__ unit_count,
                                                          Unit* __restrict__ work) {
  __shared__ uint32_t s_start[kOrderClasses];       // next free position of each class
  __shared__ uint32_t s_wcnt[16][kOrderClasses];    // units of each class per wave (a chunk)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t nu = *unit_count;
  if (tid < kOrderClasses) s_start[tid] = 0;
  __syncthreads();
  // the class sizes (order-free counts)
  for (uint32_t u = tid; u < nu; u += blockDim.x)
    atomicAdd(&s_start[order_class(units[u].count)], 1u);
  __syncthreads();
  if (tid == 0) {
    uint32_t start = 0;
    for (int c = kOrderClasses - 1; c >= 0; --c) {
      const uint32_t size = s_start[c];
      s_start[c] = start;
      start += size;
    }
  }
  __syncthreads();
  // chunks of 1024 units in table order: a unit's place in its class = the class's units of
  // the earlier chunks, of the earlier waves of this chunk and of the lower lanes of its wave
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t c0 = 0; c0 < nu; c0 += blockDim.x) {
    const uint32_t u = c0 + tid;
    const bool valid = u < nu;
    Unit w = valid ? units[u] : Unit{};
    const int c = order_class(w.count);
    uint32_t rank = 0;
#pragma unroll
    for (int k = 0; k < kOrderClasses; ++k) {
      const uint64_t m = __ballot(valid && c == k);
      if (c == k) rank = __popcll(m & below);
      if (lane == 0) s_wcnt[wave][k] = __popcll(m);
    }
    __syncthreads();
    if (valid) {
      uint32_t at = s_start[c] + rank;
      for (int v = 0; v < wave; ++v) at += s_wcnt[v][c];
      w.units = static_cast<int32_t>(u);
      work[at] = w;
    }
    __syncthreads();
    if (tid < kOrderClasses) {
      uint32_t add = 0;
      for (int v = 0; v < 16; ++v) add += s_wcnt[v][tid];
      s_start[tid] += add;
    }
    __syncthreads();
  }
}

// Sums the region sums of every brick with several units into its first unit: one
// workgroup per (such brick, 64 consecutive region entries); wave w adds units w, w+4, ...
// and the 4 partial sums meet in LDS. Integer sums: any order gives the same result.
constexpr int kReduceChunks = (kRN + 63) / 64;
constexpr int kCombineBatchR = 8;  // units in flight per lane

__global__ void __launch_bounds__(256) brick_reduce_kernel(const Unit* __restrict__ units,
                                                           const int32_t* __restrict__ multi,
                                                           const uint32_t* __restrict__ multi_count,
                                                           unsigned long long* __restrict__ partial) {
  __shared__ unsigned long long s_sum[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t items = *multi_count * kReduceChunks;
  for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int32_t u0 = multi[it / kReduceChunks];
    const int r = static_cast<int>(it % kReduceChunks) * 64 + lane;
    const int k = units[u0].units;
    unsigned long long acc = 0ull;
    if (r < kRN) {
      const unsigned long long* p = partial + static_cast<int64_t>(u0) * kRegion + r;
#pragma unroll 1
      for (int q = wave; q < k; q += 4 * kCombineBatchR) {
        unsigned long long v[kCombineBatchR];
#pragma unroll
        for (int w = 0; w < kCombineBatchR; ++w) {
          const int qq = q + 4 * w;
          v[w] = qq < k ? p[static_cast<int64_t>(qq) * kRegion] : 0ull;
        }
#pragma unroll
        for (int w = 0; w < kCombineBatchR; ++w) acc += v[w];
      }
    }
    s_sum[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && r < kRN)
      partial[static_cast<int64_t>(u0) * kRegion + r] = s_sum[0][lane] + s_sum[1][lane] +
                                                     s_sum[2][lane] + s_sum[3][lane];
    __syncthreads();
  }
}

// Node bricks list[0, count[1]) (those with sources) are stored every time; the others
// only when `full` (a new binning, another rho) or when this or the previous deposit had
// particles outside their regions (esc_any[parity], dirty[parity]), whose sums may lie
// there; otherwise they still hold the zeros stored before. The flags for the next deposit
// (and its unit pass's claim counter, zeroed) go to the other parity, which no workgroup of
// this launch reads.
//
// One wave per node brick B; lane (a, b) sums the node row (a, b, 0..7). Node brick B reads
// from the region of cell brick B + e (e in {-1, 0, 1}^3) the class block 1 - e: its own
// core, one node layer of the brick below (class 2 = node 8B) and two of the brick above
// (class 0 = nodes 8B+6, 8B+7) per axis, each block contiguous; all of a lane's loads are
// independent, so a wave has them in flight together. vec_rho: rho rows of 8 nodes are
// 16-byte aligned (n a multiple of 4, rho aligned).
__global__ void __launch_bounds__(256) brick_combine_kernel(
    const int32_t* __restrict__ node_list, uint32_t nnb, const int2* __restrict__ desc,
    const unsigned long long* __restrict__ partial, unsigned long long* __restrict__ esc,
    uint8_t* __restrict__ eflag, uint32_t* __restrict__ esc_any, uint32_t* __restrict__ dirty,
    uint32_t* __restrict__ queue, int parity, bool full, uint32_t* __restrict__ node_count,
    const unsigned int* __restrict__ maxbits, int bl, int n, int nbn,
    float* __restrict__ rho, bool vec_rho) {
  const int s = pm::fixed_scale(*maxbits, bl);
  const int lane = threadIdx.x & 63;
  const int a = lane >> 3, b = lane & 7;
  const uint32_t wave0 =
      static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6)));
  const uint32_t waves = gridDim.x * 4;
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const bool outside = esc_any[parity] != 0u;
  const uint32_t active = node_count[1];
  const uint32_t limit = (full || outside || dirty[parity]) ? nnb : active;
  for (uint32_t e = wave0; e < limit; e += waves) {
    const uint32_t B = static_cast<uint32_t>(node_list[e]);
    const bool flagged = eflag[B] != 0;
    int bi, bj, bk;
    decompose(B, static_cast<uint32_t>(nbn), bi, bj, bk);
    const int i = kCB * bi + a, j = kCB * bj + b, k0 = kCB * bk;
    unsigned long long sum[kCB];
#pragma unroll
    for (int c = 0; c < kCB; ++c) sum[c] = 0ull;
    if (e < active) {
#pragma unroll
      for (int d = 0; d < 27; ++d) {
        const int ex = d / 9 - 1, ey = (d / 3) % 3 - 1, ez = d % 3 - 1;
        const int2 dd = desc[static_cast<int64_t>(e) * 27 + d];
        // the lane's row lies in this block when its a (b) is a layer the block covers
        const bool row = (ex == 0 || (ex < 0 ? a == 0 : a >= 6)) &&
                         (ey == 0 || (ey < 0 ? b == 0 : b >= 6));
        if (dd.y > 0 && row) {
          const int ny = cls_n(1 - ey), nz = cls_n(1 - ez);
          const int ox = ex > 0 ? a - 6 : (ex == 0 ? a : 0);
          const int oy = ey > 0 ? b - 6 : (ey == 0 ? b : 0);
          // rows of 8 and of 2 start at even entries of 16-byte aligned regions
          const unsigned long long* src = partial + static_cast<int64_t>(dd.x) * kRegion +
                                          block_off(1 - ex, 1 - ey, 1 - ez) + (ox * ny + oy) * nz;
          // nodes c of the row in this block: all 8 (ez = 0), c = 0 (ez = -1), c = 6, 7
          if (ez == 0) {
#pragma unroll
            for (int c = 0; c < kCB; c += 2) {
              const ulonglong2 w = *reinterpret_cast<const ulonglong2*>(src + c);
              sum[c] += w.x;
              sum[c + 1] += w.y;
            }
          } else if (ez < 0) {
            sum[0] += src[0];
          } else {
            const ulonglong2 w = *reinterpret_cast<const ulonglong2*>(src);
            sum[6] += w.x;
            sum[7] += w.y;
          }
        }
      }
    }
    const bool in_ij = i < n && j < n;
    if (flagged && in_ij) {
#pragma unroll
      for (int c = 0; c < kCB; ++c) {
        if (k0 + c < n) {
          const int64_t nd = i * n2 + static_cast<int64_t>(j) * n + (k0 + c);
          sum[c] += esc[nd];
          esc[nd] = 0ull;
        }
      }
    }
    if (in_ij) {
      float* dst = rho + i * n2 + static_cast<int64_t>(j) * n + k0;
      float r[kCB];
#pragma unroll
      for (int c = 0; c < kCB; ++c)
        r[c] = static_cast<float>(ldexp(static_cast<double>(sum[c]), -s));
      if (vec_rho && k0 + kCB <= n) {
        reinterpret_cast<float4*>(dst)[0] = make_float4(r[0], r[1], r[2], r[3]);
        reinterpret_cast<float4*>(dst)[1] = make_float4(r[4], r[5], r[6], r[7]);
      } else {
#pragma unroll
        for (int c = 0; c < kCB; ++c)
          if (k0 + c < n) dst[c] = r[c];
      }
    }
    if (flagged && lane == 0) eflag[B] = 0;  // every lane read it above
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the gradient covers every node brick when a particle deposited outside the regions
    node_count[0] = outside ? nnb : active;
    dirty[parity ^ 1] = outside ? 1u : 0u;
    esc_any[parity ^ 1] = 0u;
    queue[parity ^ 1] = 0u;  // the next deposit's unit pass claims from it
  }
}

__global__ void __launch_bounds__(256) zero_fill_kernel(float* __restrict__ p, int64_t count,
                                                        uint32_t* __restrict__ node_count) {
  const int64_t i = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0.0f;
  if (i == 0) node_count[0] = node_count[1] = 0u;
}

template <typename T>
pm_status alloc(T** ptr, int64_t count, const char* what) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), sizeof(T) * (count > 0 ? count : 1));
  if (e != hipSuccess) return pm::hip_fail(what, e);
  return PM_OK;
}

}  // namespace

namespace pm {
namespace depbrick {

struct State {
  int n = 0, nbc = 0, nbn = 0;
  int64_t cap = 0, ncb = 0, nnb = 0, max_units = 0, max_home_units = 0;
  int passes = 0;
  int bits[4] = {0, 0, 0, 0};
  int32_t* first = nullptr;
  int32_t* end = nullptr;
  int32_t* unit_first = nullptr;
  int32_t* unit_num = nullptr;
  Unit* units = nullptr;
  uint32_t* unit_count = nullptr;
  uint32_t* node_src = nullptr;
  int32_t* node_list = nullptr;
  uint32_t* node_count = nullptr;  // [0] gradient count, [1] node bricks with sources
  int2* desc = nullptr;            // 27 source descriptors per node brick with sources
  int32_t* multi = nullptr;        // first units of the bricks with several units
  uint32_t* multi_count = nullptr;
  Unit* work = nullptr;  // the units in the order the unit pass claims them (4th field: index)
  uint32_t* queue = nullptr;       // claim counters of the unit pass, by parity
  unsigned long long* partial = nullptr;
  unsigned long long* esc = nullptr;
  uint8_t* eflag = nullptr;
  uint32_t* esc_any = nullptr;  // [2]: particles outside their regions, by deposit parity
  uint32_t* dirty = nullptr;    // [2]: the deposit before had such particles
  unsigned int* maxbits = nullptr;
  // host side
  bool valid = false;     // a binning exists
  int64_t np = -1;        // particles it covers
  float dx = 0.0f;
  const int32_t* perm = nullptr;  // its order (in the plan's sort buffers)
  bool ordered = false;   // the particles lie in that order
  bool binned_now = false;
  int parity = 0;         // flips every deposit
  const float* last_rho = nullptr;  // rho of the last deposit (all of it written then)
};

void destroy(State* st) {
  if (!st) return;
  void* bufs[] = {st->first, st->end, st->unit_first, st->unit_num, st->units, st->unit_count,
                  st->node_src, st->node_list, st->node_count, st->desc, st->multi,
                  st->multi_count, st->work, st->queue, st->partial, st->esc, st->eflag,
                  st->esc_any, st->dirty, st->maxbits};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete st;
}

pm_status create(int n, int64_t cap, State** out) {
  *out = nullptr;
  // the unit pass addresses particles by 32-bit byte offsets
  if (cap >= (int64_t(1) << 30))
    return pm::invalid("PM_DEPOSIT_BRICKED: at most 2^30 - 1 particles per plan");
  auto* st = new State();
  st->n = n;
  st->cap = cap;
  st->nbc = (n + 1 + kCB - 1) / kCB;
  st->nbn = (n + kCB - 1) / kCB;
  st->ncb = static_cast<int64_t>(st->nbc) * st->nbc * st->nbc;
  st->nnb = static_cast<int64_t>(st->nbn) * st->nbn * st->nbn;
  const int64_t full = (cap + kUnit - 1) / kUnit;
  st->max_home_units = full + std::min<int64_t>(st->ncb, cap);
  st->max_units = st->max_home_units + full + 1;
  st->passes = pm::radix::plan_digits(static_cast<uint64_t>(st->ncb) << kCellBits, st->bits);
  const int64_t nodes = static_cast<int64_t>(n) * n * n;
  pm_status e = PM_OK;
  // the one-workgroup table passes read whole rounds: first/end/node_src are padded
  if ((e = alloc(&st->first, padded(st->ncb), "hipMalloc (brick bounds)")) ||
      (e = alloc(&st->end, padded(st->ncb), "hipMalloc (brick bounds)")) ||
      (e = alloc(&st->unit_first, st->ncb, "hipMalloc (brick units)")) ||
      (e = alloc(&st->unit_num, st->ncb, "hipMalloc (brick units)")) ||
      (e = alloc(&st->units, st->max_units, "hipMalloc (units)")) ||
      (e = alloc(&st->unit_count, 1, "hipMalloc (unit count)")) ||
      (e = alloc(&st->node_src, padded(st->nnb), "hipMalloc (node sources)")) ||
      (e = alloc(&st->node_list, st->nnb, "hipMalloc (node list)")) ||
      (e = alloc(&st->node_count, 2, "hipMalloc (node count)")) ||
      (e = alloc(&st->desc, 27 * st->nnb, "hipMalloc (node sources)")) ||
      (e = alloc(&st->multi, st->max_units, "hipMalloc (brick units)")) ||
      (e = alloc(&st->multi_count, 1, "hipMalloc (brick units)")) ||
      (e = alloc(&st->work, st->max_units, "hipMalloc (unit order)")) ||
      (e = alloc(&st->queue, 2, "hipMalloc (unit queue)")) ||
      (e = alloc(&st->partial, st->max_home_units * kRegion, "hipMalloc (unit sums)")) ||
      (e = alloc(&st->esc, nodes, "hipMalloc (outside sums)")) ||
      (e = alloc(&st->eflag, st->nnb, "hipMalloc (outside flags)")) ||
      (e = alloc(&st->esc_any, 2, "hipMalloc (outside flag)")) ||
      (e = alloc(&st->dirty, 2, "hipMalloc (outside flag)")) ||
      (e = alloc(&st->maxbits, 1 + pm::depfix::kMaxBlocks, "hipMalloc (deposit scale)"))) {
    destroy(st);
    return e;
  }
  // zero state the kernels keep zeroed behind themselves (set up once, synchronously)
  hipError_t he = hipSuccess;
  if ((he = hipMemset(st->first, 0, sizeof(int32_t) * padded(st->ncb))) != hipSuccess ||
      (he = hipMemset(st->end, 0, sizeof(int32_t) * padded(st->ncb))) != hipSuccess ||
      (he = hipMemset(st->node_src, 0, sizeof(uint32_t) * padded(st->nnb))) != hipSuccess ||
      (he = hipMemset(st->esc, 0, sizeof(unsigned long long) * std::max<int64_t>(nodes, 1))) !=
          hipSuccess ||
      (he = hipMemset(st->eflag, 0, st->nnb)) != hipSuccess ||
      (he = hipMemset(st->esc_any, 0, 2 * sizeof(uint32_t))) != hipSuccess ||
      (he = hipMemset(st->dirty, 0, 2 * sizeof(uint32_t))) != hipSuccess ||
      (he = hipMemset(st->queue, 0, 2 * sizeof(uint32_t))) != hipSuccess ||

      (he = hipMemset(st->node_count, 0, 2 * sizeof(uint32_t))) != hipSuccess ||
      (he = hipDeviceSynchronize()) != hipSuccess) {
    destroy(st);
    return pm::hip_fail("hipMemset (bricked deposit state)", he);
  }
  *out = st;
  return PM_OK;
}

pm_status deposit(State* st, const SortBuffers& sb, const float* px, const float* py,
                  const float* pz, const float* mass, int64_t np, float dx, float* rho, bool bin,
                  bool rho_kept, hipStream_t s) {
  PM_REQUIRE(np >= 0 && np <= st->cap, "bricked deposit: particle count out of range");
  const int n = st->n;
  const int64_t nodes = static_cast<int64_t>(n) * n * n;
  st->binned_now = false;
  if (np == 0) {
    zero_fill_kernel<<<pm::blocks_for(nodes), 256, 0, s>>>(rho, nodes, st->node_count);
    PM_CHECK_LAUNCH("zero_fill_kernel");
    st->valid = false;
    st->last_rho = nullptr;
    return PM_OK;
  }
  const float inv_vol = 1.0f / ((dx * dx) * dx);
  const int bl = pm::bit_length(np);
  if (bin || !st->valid || st->np != np || st->dx != dx) {
    pm_status e;
    const unsigned key_blocks =
        static_cast<unsigned>(std::min<int64_t>(pm::blocks_for(np), pm::depfix::kMaxBlocks));
    {
      PM_TIMED("deposit.brick_key", 20.0 * np, s);
      brick_key_kernel<<<key_blocks, 256, 0, s>>>(
          px, py, pz, mass, np, n, dx, inv_vol, static_cast<uint32_t>(st->nbc),
          static_cast<uint32_t>(st->ncb) << kCellBits, sb.keys[0], st->maxbits + 1);
      PM_CHECK_LAUNCH("brick_key_kernel");
    }
    int src = 0;
    if ((e = pm::radix::sort(sb.keys[0], sb.keys, sb.vals, sb.hist, sb.digit, np, 0, st->bits,
                             st->passes, &src, s)))
      return e;
    {
      PM_TIMED("deposit.bin_bounds", 4.0 * np, s);
      brick_bounds_kernel<<<pm::blocks_for(np), 256, 0, s>>>(
          sb.keys[src], np, static_cast<uint32_t>(st->ncb), st->first, st->end);
      PM_CHECK_LAUNCH("brick_bounds_kernel");
    }
    {
      PM_TIMED("deposit.bin_units", 0.0, s);
      unit_build_kernel<<<1, 1024, 0, s>>>(st->first, st->end, static_cast<uint32_t>(st->ncb),
                                           np, kUnit, st->units, st->unit_first, st->unit_num,
                                           st->unit_count, st->maxbits,
                                           static_cast<int>(key_blocks), st->multi,
                                           st->multi_count);
      PM_CHECK_LAUNCH("unit_build_kernel");
      unit_order_kernel<<<1, 1024, 0, s>>>(st->units, st->unit_count, st->work);
      PM_CHECK_LAUNCH("unit_order_kernel");
    }
    {
      PM_TIMED("deposit.bin_nodes", 0.0, s);
      node_src_kernel<<<pm::blocks_for(st->nnb), 256, 0, s>>>(st->unit_num, st->nbc, st->nbn,
                                                               st->node_src);
      PM_CHECK_LAUNCH("node_src_kernel");
      node_list_kernel<<<1, 1024, 0, s>>>(st->node_src, st->nnb, st->node_list, st->node_count);
      PM_CHECK_LAUNCH("node_list_kernel");
      node_desc_kernel<<<pm::blocks_for(27 * st->nnb), 256, 0, s>>>(
          st->node_list, st->node_count, st->node_src, st->unit_first, st->unit_num, st->nbc,
          st->nbn, st->desc);
      PM_CHECK_LAUNCH("node_desc_kernel");
    }
    st->perm = sb.vals[src];
    st->ordered = false;
    st->valid = true;
    st->np = np;
    st->dx = dx;
    st->binned_now = true;
  }
  const int32_t* idx = st->ordered ? nullptr : st->perm;
  {
    const double bytes = (idx ? 20.0 : 16.0) * np;
#define PM_UNIT_PASS(F64, IDX)                                                                 \
  PM_LAUNCH_TIMED("deposit.unit_sums", bytes, (unit_deposit_kernel<F64, IDX>), kUnitGroups, 256, \
                  0, s, st->work, st->unit_count, st->queue, idx, px, py, pz, mass,            \
                  np, n, dx, inv_vol,                                                          \
                  st->nbc, st->nbn, st->maxbits, bl, st->partial, st->esc, st->eflag,          \
                  st->esc_any, st->parity)
    if (bl >= kF64Bits) {
      if (idx) PM_UNIT_PASS(true, true); else PM_UNIT_PASS(true, false);
    } else {
      if (idx) PM_UNIT_PASS(false, true); else PM_UNIT_PASS(false, false);
    }
#undef PM_UNIT_PASS
    PM_CHECK_LAUNCH("unit_deposit_kernel");
  }
  {
    PM_LAUNCH_TIMED("deposit.brick_reduce", 0.0, brick_reduce_kernel, kPersistent, 256, 0, s,
                    st->units, st->multi, st->multi_count, st->partial);
    PM_CHECK_LAUNCH("brick_reduce_kernel");
  }
  {
    const bool full = st->binned_now || !rho_kept || rho != st->last_rho;
    PM_LAUNCH_TIMED("deposit.combine", 4.0 * nodes, brick_combine_kernel,
                    static_cast<unsigned>(std::min<int64_t>(st->nnb, kPersistent)), 256, 0, s,
                    st->node_list, static_cast<uint32_t>(st->nnb), st->desc, st->partial, st->esc,
                    st->eflag, st->esc_any, st->dirty, st->queue, st->parity, full,
                    st->node_count, st->maxbits, bl, n, st->nbn, rho,
                    (n & 3) == 0 && (reinterpret_cast<uintptr_t>(rho) & 15u) == 0u);
    PM_CHECK_LAUNCH("brick_combine_kernel");
  }
  st->parity ^= 1;
  st->last_rho = rho;
  return PM_OK;
}

pm_status mark_ordered(State* st) {
  PM_REQUIRE(st->valid, "pm_deposit_plan_mark_ordered: no binning to be in the order of");
  st->ordered = true;
  return PM_OK;
}

const int32_t* order(const State* st) { return st->binned_now ? st->perm : nullptr; }

void node_list(const State* st, const int32_t** list, const uint32_t** count, int64_t* max) {
  // the count is that of the last deposit (0 before the first)
  *list = st->node_list;
  *count = st->node_count;
  if (max) *max = st->nnb;
}

void invalidate(State* st) {
  st->valid = false;
  st->binned_now = false;
}

}  // namespace depbrick
}  // namespace pm
