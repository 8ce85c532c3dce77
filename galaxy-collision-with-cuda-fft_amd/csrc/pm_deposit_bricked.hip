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
// Binning (when asked for, or when there is none): key = cell brick (or a sentinel past
// the last brick for particles with no corner in the box), a stable radix sort of the keys,
// brick bounds in the sorted order, and a unit table: each brick's positions cut into units
// of at most kUnit, then the positions of the sentinel particles as units of no brick.
//
// Every deposit (two launches):
//   unit pass: one workgroup per unit zeroes the 1331 LDS sums, reads its particles (through
//     the binning's order, or in place once the caller has moved the particles into that
//     order), adds the 8 corner q of every particle whose cell lies in the region with LDS
//     atomics (a wave whose 64 particles share one cell sums them with shuffles first and
//     adds once), and stores the region's sums for its unit. A particle outside the
//     region (it moved more than one cell out of its brick since the binning, or has no
//     brick) adds its q to a zeroed n^3 64-bit array with global atomics and flags the node
//     bricks it touched: exact, only slower, so the result never depends on how fresh the
//     binning is;
//   combine pass: every 8^3 node brick sums, per node, the unit sums of the cell bricks whose
//     region holds the node (the brick itself; the one below for the first layer; the one
//     above for the last two), adds and clears the flagged 64-bit array entries, converts,
//     and stores rho; node bricks with no unit nearby store zeros.
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
constexpr int kR = kCB + 1 + 2;          // region nodes per axis (one cell of slack a side)
constexpr int kRN = kR * kR * kR;        // 1331 sums per region
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
  int32_t units;  // first unit of a brick: the brick's number of units; others 0
};

__device__ __forceinline__ void decompose(uint32_t b, uint32_t nb, int& x, int& y, int& z) {
  z = static_cast<int>(b % nb);
  y = static_cast<int>((b / nb) % nb);
  x = static_cast<int>(b / (nb * nb));
}

// ---------------------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------------------
// Cell brick of every particle (grid-stride), and this workgroup's largest finite positive
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
    keys[p] = in ? (bx * nbc + by) * nbc + bz : sentinel;
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
  const uint32_t b = skeys[j];
  const uint32_t prev = j > 0 ? skeys[j - 1] : 0xFFFFFFFFu;
  const uint32_t next = j + 1 < np ? skeys[j + 1] : 0xFFFFFFFFu;
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

// Per-thread span of the one-workgroup table passes: 32 consecutive entries read as eight
// 16-byte loads issued together (the tables are padded to whole rounds of 1024 spans).
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
    const uint32_t b0 = round + threadIdx.x * kSpan;
    int4 f4[kSpan / 4], e4[kSpan / 4];
#pragma unroll
    for (int q = 0; q < kSpan / 4; ++q) {
      f4[q] = reinterpret_cast<const int4*>(first + b0)[q];
      e4[q] = reinterpret_cast<const int4*>(end + b0)[q];
    }
    int32_t f[kSpan], c[kSpan];
#pragma unroll
    for (int q = 0; q < kSpan / 4; ++q) {
      f[4 * q] = f4[q].x, f[4 * q + 1] = f4[q].y, f[4 * q + 2] = f4[q].z, f[4 * q + 3] = f4[q].w;
      c[4 * q] = e4[q].x - f4[q].x, c[4 * q + 1] = e4[q].y - f4[q].y;
      c[4 * q + 2] = e4[q].z - f4[q].z, c[4 * q + 3] = e4[q].w - f4[q].w;
    }
    uint32_t nu = 0, nc = 0, nm = 0;
#pragma unroll
    for (int q = 0; q < kSpan; ++q) {
      const uint32_t k = (static_cast<uint32_t>(c[q]) + unit - 1) / unit;
      nu += k;
      nc += static_cast<uint32_t>(c[q]);
      nm += k >= 2 ? 1u : 0u;
    }
    uint32_t round_units, round_binned, round_multi;
    uint32_t base = units_before + scan_1024(nu, round_units, s_wave);
    (void)scan_1024(nc, round_binned, s_wave);
    uint32_t mbase = multi_before + scan_1024(nm, round_multi, s_wave);
#pragma unroll
    for (int q = 0; q < kSpan; ++q) {
      const uint32_t b = b0 + q;
      if (b >= ncb) break;
      const uint32_t cnt = static_cast<uint32_t>(c[q]);
      const uint32_t k = (cnt + unit - 1) / unit;
      unit_first[b] = static_cast<int32_t>(base);
      unit_num[b] = static_cast<int32_t>(k);
      for (uint32_t w = 0; w < k; ++w) {
        Unit un;
        un.brick = static_cast<int32_t>(b);
        un.begin = f[q] + static_cast<int32_t>(w * unit);
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
    const int64_t lo = round + int64_t(threadIdx.x) * kSpan;
    uint4 m4[kSpan / 4];
#pragma unroll
    for (int q = 0; q < kSpan / 4; ++q) m4[q] = reinterpret_cast<const uint4*>(node_src + lo)[q];
    uint32_t bits = 0, act = 0, idle = 0;
#pragma unroll
    for (int q = 0; q < kSpan; ++q) {
      const uint4 w = m4[q / 4];
      const uint32_t v = (q & 3) == 0 ? w.x : (q & 3) == 1 ? w.y : (q & 3) == 2 ? w.z : w.w;
      const bool valid = lo + q < nnb;
      bits |= (valid && v) ? (1u << q) : 0u;
      act += (valid && v) ? 1u : 0u;
      idle += (valid && !v) ? 1u : 0u;
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
    for (int q = 0; q < kSpan; ++q) {
      const int64_t B = lo + q;
      if (B >= nnb) break;
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
  // t * 2^s is exact in float32 (a power-of-two scaling that stays far from overflow and
  // from the subnormals), so rounding it there gives llrint(ldexp((double)t, s))
  static __device__ __forceinline__ T term(float t, int s) {
    return static_cast<double>(rintf(__builtin_amdgcn_ldexpf(t, s)));
  }
  static __device__ __forceinline__ unsigned long long u64(T v) {
    return static_cast<unsigned long long>(v);
  }
};
template <>
struct Sum<false> {
  using T = unsigned long long;
  static __device__ __forceinline__ T term(float t, int s) { return pm::fixed_q(t, s); }
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

// Adds the 8 corner sums of the lanes with `flushing` set to their cell `cur`. Lanes walk
// the binned positions side by side, so where a wave crosses from one cell to the next many
// lanes flush the same cell at once: up to two such groups of 16 or more lanes are summed
// over the wave first (one add per row and corner), the rest add per lane. Every lane of
// the wave must call this.
template <typename T>
__device__ __forceinline__ void flush_sums(T* __restrict__ acc, int cur, const T (&sum)[8],
                                           bool flushing, int lane) {
  uint64_t fl = __ballot(flushing);
#pragma unroll 1
  for (int it = 0; it < 2 && __popcll(fl) >= 16; ++it) {
    const int ka = __builtin_amdgcn_readlane(cur, __builtin_ctzll(fl));
    const int kz = __builtin_amdgcn_readlane(cur, 63 - __builtin_clzll(fl));
    const uint64_t ma = __ballot(cur == ka) & fl, mz = __ballot(cur == kz) & fl;
    const int k0 = __popcll(mz) > __popcll(ma) ? kz : ka;
    const uint64_t same = __popcll(mz) > __popcll(ma) ? mz : ma;
    if (__popcll(same) < 16) break;
    const bool me = (same >> lane) & 1ull;
    T w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = me ? sum[c] : T(0);
    T z = group_sums(w, lane);
    z = z + dpp64<kDppRor8>(z);
    if ((lane & 8) == 0) atomicAdd(&acc[k0 + corner_off(lane & 7)], z);
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

// Each lane walks kRun binned positions 64 apart (wave w's positions of a round are
// w*kWavePos + 64*i + lane, so every load is coalesced; cell-ordered particles keep a lane
// in one cell for many steps where cells are dense, with the few that drifted into a
// neighbouring cell since the binning in between) and keeps the 8 corner sums of up to
// three cells in registers; they go to the LDS sums at the end of the round, those that
// many lanes share summed over the wave first (flush_sums).
constexpr int kRun = 16;                  // positions per lane and round
constexpr int kWavePos = 64 * kRun;       // positions per wave and round
constexpr int kRoundPos = 4 * kWavePos;   // positions per workgroup and round
constexpr int kGroup = 4;                 // particles per lane whose loads are in flight

// IDX: the particles are read through the binning's order idx (else in place).
// Workgroup w takes units [wg_first[w], wg_first[w+1]) (unit_assign_kernel).
template <bool F64, bool IDX>
__global__ void __launch_bounds__(256) unit_deposit_kernel(
    const Unit* __restrict__ units, const int32_t* __restrict__ wg_first,
    const int32_t* __restrict__ idx, const float* __restrict__ px, const float* __restrict__ py,
    const float* __restrict__ pz, const float* __restrict__ mass, int n, float dx,
    float inv_vol, int nbc, int nbn, const unsigned int* __restrict__ maxbits, int bl,
    unsigned long long* __restrict__ partial, unsigned long long* __restrict__ esc,
    uint8_t* __restrict__ eflag, uint32_t* __restrict__ esc_any, int parity) {
  using S = Sum<F64>;
  using T = typename S::T;
  __shared__ T acc[kRN];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int s = pm::fixed_scale(*maxbits, bl);
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int32_t u_end = wg_first[blockIdx.x + 1];
  for (int32_t u = wg_first[blockIdx.x]; u < u_end; ++u) {
    const Unit un = units[u];
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
    // a group's particles (positions past the unit's end read its last particle, unused);
    // the next group's loads are issued before this group's work
    auto load = [&](int off0, float4 (&p)[kGroup]) {
      int32_t sl[kGroup];
#pragma unroll
      for (int i = 0; i < kGroup; ++i) {
        const int32_t j = un.begin + min(off0 + 64 * i, un.count - 1);
        sl[i] = IDX ? idx[j] : j;
      }
#pragma unroll
      for (int i = 0; i < kGroup; ++i)
        p[i] = make_float4(px[sl[i]], py[sl[i]], pz[sl[i]], mass[sl[i]]);
    };
    for (int base = 0; base < un.count; base += kRoundPos) {
      const int p0 = base + wave * kWavePos + lane;
      // three cells (region indices, -1 when free) with their 8 corner sums: the lane's
      // cell, and the cells of particles that drifted into a neighbouring cell or lie
      // past a cell boundary; a fourth cell takes slot 2 after adding its sums
      int k0 = -1, k1 = -1, k2 = -1;
      T s0[8], s1[8], s2[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) s0[c] = s1[c] = s2[c] = T(0);
      float4 grp[kGroup];
      load(p0, grp);
      // (not unrolled: the unrolled body of all kRun particles outgrew the instruction cache)
#pragma unroll 1
      for (int g = 0; g < kRun; g += kGroup) {
        float4 nxt[kGroup];
        if (g + kGroup < kRun) load(p0 + 64 * (g + kGroup), nxt);
#pragma unroll
        for (int i = 0; i < kGroup; ++i) {
          const bool valid = p0 + 64 * (g + i) < un.count;
          int ix = 0, iy = 0, iz = 0;
          float fx = 0.0f, fy = 0.0f, fz = 0.0f;
          const bool in = valid && pm::cic_axis(grp[i].x, dx, n, ix, fx) &&
                          pm::cic_axis(grp[i].y, dx, n, iy, fy) &&
                          pm::cic_axis(grp[i].z, dx, n, iz, fz);
          const float mp = grp[i].w * inv_vol;
          // m' <= 0 or not finite: every term is dropped (t <= 0 or not finite)
          const bool good = in && mp > 0.0f && mp < __builtin_huge_valf();
          const int rx = ix - ox, ry = iy - oy, rz = iz - oz;
          const bool reg = good && home && static_cast<unsigned>(rx) <= 9u &&
                           static_cast<unsigned>(ry) <= 9u && static_cast<unsigned>(rz) <= 9u;
          if (!good) continue;
          // t_c = ((m'*wx)*wy)*wz, the products shared between corners
          const float gx0 = 1.0f - fx, gy0 = 1.0f - fy, gz0 = 1.0f - fz;
          const float a0 = mp * gx0, a1 = mp * fx;
          const float bxy[4] = {a0 * gy0, a0 * fy, a1 * gy0, a1 * fy};
          T v[8];
#pragma unroll
          for (int c = 0; c < 8; ++c) v[c] = S::term(bxy[c >> 1] * ((c & 1) ? fz : gz0), s);
          if (reg) {
            const int key = (rx * kR + ry) * kR + rz;
            if (key == k0 || k0 < 0) {
              k0 = key;
#pragma unroll
              for (int c = 0; c < 8; ++c) s0[c] = s0[c] + v[c];
            } else if (key == k1 || k1 < 0) {
              k1 = key;
#pragma unroll
              for (int c = 0; c < 8; ++c) s1[c] = s1[c] + v[c];
            } else if (key == k2 || k2 < 0) {
              k2 = key;
#pragma unroll
              for (int c = 0; c < 8; ++c) s2[c] = s2[c] + v[c];
            } else {
#pragma unroll
              for (int c = 0; c < 8; ++c) {
                atomicAdd(&acc[k2 + corner_off(c)], s2[c]);
                s2[c] = v[c];
              }
              k2 = key;
            }
          } else {
            unsigned long long q[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) q[c] = S::u64(v[c]);
            add_outside(q, ix, iy, iz, n, nbn, esc, eflag, esc_any + parity);
          }
        }
        if (g + kGroup < kRun) {
#pragma unroll
          for (int i = 0; i < kGroup; ++i) grp[i] = nxt[i];
        }
      }
      flush_sums(acc, k0, s0, k0 >= 0, lane);
      flush_sums(acc, k1, s1, k1 >= 0, lane);
      flush_sums(acc, k2, s2, k2 >= 0, lane);
    }
    __syncthreads();
    if (home) {
      unsigned long long* dst = partial + static_cast<int64_t>(u) * kRN;
      for (int q = tid; q < kRN; q += blockDim.x) dst[q] = S::u64(acc[q]);
    }
    __syncthreads();  // the next unit zeroes acc
  }
}

// One workgroup: splits the units into `groups` consecutive ranges of about equal cost
// (positions plus kUnitCost per unit, for its setup and its stored region), so the unit
// pass's workgroups finish together: wg_first[w] = the first unit of range w,
// wg_first[groups] = the number of units.
constexpr int kUnitCost = 1024;

__global__ void __launch_bounds__(1024) unit_assign_kernel(const Unit* __restrict__ units,
                                                           const uint32_t* __restrict__ unit_count,
                                                           int groups,
                                                           int32_t* __restrict__ wg_first) {
  __shared__ uint32_t s_wave[16];
  const uint32_t nu = *unit_count;
  const uint32_t span = (nu + 1023u) / 1024u;
  const uint32_t u0 = min(threadIdx.x * span, nu), u1 = min(u0 + span, nu);
  uint32_t cost = 0;
  for (uint32_t u = u0; u < u1; ++u) cost += static_cast<uint32_t>(units[u].count) + kUnitCost;
  uint32_t total;
  uint32_t before = scan_1024(cost, total, s_wave);
  // unit u starts range floor(P_u * groups / total) (P_u: cost of the units before u); the
  // ranges between the previous unit's and u's take u as their first unit
  int64_t prev_w = -1;
  if (u0 > 0 && u0 < u1) {
    const uint32_t pc = static_cast<uint32_t>(units[u0 - 1].count) + kUnitCost;
    prev_w = static_cast<int64_t>(before - pc) * groups / max(total, 1u);
  }
  for (uint32_t u = u0; u < u1; ++u) {
    const int64_t w = static_cast<int64_t>(before) * groups / max(total, 1u);
    for (int64_t q = prev_w + 1; q <= w; ++q) wg_first[q] = static_cast<int32_t>(u);
    prev_w = w;
    before += static_cast<uint32_t>(units[u].count) + kUnitCost;
  }
  __syncthreads();
  // ranges after the last unit's (and all of them when there are no units) are empty
  if (threadIdx.x == 0) {
    int64_t last_w = -1;
    if (nu > 0) {
      const uint32_t pc = static_cast<uint32_t>(units[nu - 1].count) + kUnitCost;
      last_w = static_cast<int64_t>(total - pc) * groups / max(total, 1u);
    }
    for (int64_t q = last_w + 1; q <= groups; ++q) wg_first[q] = static_cast<int32_t>(nu);
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
      const unsigned long long* p = partial + static_cast<int64_t>(u0) * kRN + r;
#pragma unroll 1
      for (int q = wave; q < k; q += 4 * kCombineBatchR) {
        unsigned long long v[kCombineBatchR];
#pragma unroll
        for (int w = 0; w < kCombineBatchR; ++w) {
          const int qq = q + 4 * w;
          v[w] = qq < k ? p[static_cast<int64_t>(qq) * kRN] : 0ull;
        }
#pragma unroll
        for (int w = 0; w < kCombineBatchR; ++w) acc += v[w];
      }
    }
    s_sum[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && r < kRN)
      partial[static_cast<int64_t>(u0) * kRN + r] = s_sum[0][lane] + s_sum[1][lane] +
                                                     s_sum[2][lane] + s_sum[3][lane];
    __syncthreads();
  }
}

// Node bricks list[0, count[1]) (those with sources) are stored every time; the others
// only when `full` (a new binning, another rho) or when this or the previous deposit had
// particles outside their regions (esc_any[parity], dirty[parity]), whose sums may lie
// there; otherwise they still hold the zeros stored before. The flags for the next deposit
// go to the other parity, which no workgroup of this launch reads.
__global__ void __launch_bounds__(256) brick_combine_kernel(
    const int32_t* __restrict__ node_list, uint32_t nnb, const int2* __restrict__ desc,
    const unsigned long long* __restrict__ partial, unsigned long long* __restrict__ esc,
    uint8_t* __restrict__ eflag, uint32_t* __restrict__ esc_any, uint32_t* __restrict__ dirty,
    int parity, bool full, uint32_t* __restrict__ node_count,
    const unsigned int* __restrict__ maxbits, int bl, int n, int nbn,
    float* __restrict__ rho) {
  __shared__ int2 s_desc[27];
  const int s = pm::fixed_scale(*maxbits, bl);
  const int tid = threadIdx.x;
  // lane: node row (a, bb) of the brick and nodes 2h, 2h+1 of the row
  const int a = tid >> 5, bb = (tid >> 2) & 7, h = tid & 3;
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const bool outside = esc_any[parity] != 0u;
  const uint32_t active = node_count[1];
  const uint32_t limit = (full || outside || dirty[parity]) ? nnb : active;
  for (uint32_t e = blockIdx.x; e < limit; e += gridDim.x) {
    const uint32_t B = static_cast<uint32_t>(node_list[e]);
    const bool sourced = e < active;
    if (sourced && tid < 27) s_desc[tid] = desc[static_cast<int64_t>(e) * 27 + tid];
    const bool flagged = eflag[B] != 0;
    int bi, bj, bk;
    decompose(B, static_cast<uint32_t>(nbn), bi, bj, bk);
    const int i = kCB * bi + a, j = kCB * bj + bb, k = kCB * bk + 2 * h;
    __syncthreads();
    // a node takes sums from the brick itself and, on the brick's first or last layers
    // along an axis, from the neighbour below or above: at most 8 sources, whose loads all
    // leave before any is added
    unsigned long long sum0 = 0ull, sum1 = 0ull;
    if (sourced) {
      const int ddx = a == 0 ? -1 : (a >= 6 ? 1 : 0);
      const int ddy = bb == 0 ? -1 : (bb >= 6 ? 1 : 0);
      unsigned long long v[2][8];
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int l = 2 * h + o;
        const int ddz = l == 0 ? -1 : (l >= 6 ? 1 : 0);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int ex = (c >> 2) ? ddx : 0, ey = ((c >> 1) & 1) ? ddy : 0, ez = (c & 1) ? ddz : 0;
          // the secondary source exists only where the neighbour's region reaches
          const bool used = ((c >> 2) == 0 || ddx != 0) && (((c >> 1) & 1) == 0 || ddy != 0) &&
                            ((c & 1) == 0 || ddz != 0);
          const int2 dd = s_desc[(ex + 1) * 9 + (ey + 1) * 3 + (ez + 1)];
          const int r = ((a + 2 - kCB * ex) * kR + (bb + 2 - kCB * ey)) * kR + (l + 2 - kCB * ez);
          v[o][c] = (used && dd.y > 0) ? partial[static_cast<int64_t>(dd.x) * kRN + r] : 0ull;
        }
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        sum0 += v[0][c];
        sum1 += v[1][c];
      }
    }
    const bool in_ij = i < n && j < n;
    unsigned long long sum[2] = {sum0, sum1};
    if (flagged) {
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        if (in_ij && k + o < n) {
          const int64_t node = i * n2 + static_cast<int64_t>(j) * n + (k + o);
          sum[o] += esc[node];
          esc[node] = 0ull;
        }
      }
    }
    if (in_ij) {
      float* dst = rho + i * n2 + static_cast<int64_t>(j) * n + k;
#pragma unroll
      for (int o = 0; o < 2; ++o)
        if (k + o < n) dst[o] = static_cast<float>(ldexp(static_cast<double>(sum[o]), -s));
    }
    __syncthreads();  // s_desc and the flag are read by every lane
    if (flagged && tid == 0) eflag[B] = 0;
  }
  if (blockIdx.x == 0 && tid == 0) {
    // the gradient covers every node brick when a particle deposited outside the regions
    node_count[0] = outside ? nnb : active;
    dirty[parity ^ 1] = outside ? 1u : 0u;
    esc_any[parity ^ 1] = 0u;
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
  int32_t* wg_first = nullptr;     // unit ranges of the unit pass's workgroups
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
                  st->multi_count, st->wg_first, st->partial, st->esc, st->eflag,
                  st->esc_any, st->dirty, st->maxbits};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete st;
}

pm_status create(int n, int64_t cap, State** out) {
  *out = nullptr;
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
  st->passes = pm::radix::plan_digits(static_cast<uint64_t>(st->ncb), st->bits);
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
      (e = alloc(&st->wg_first, kUnitGroups + 1, "hipMalloc (unit ranges)")) ||
      (e = alloc(&st->partial, st->max_home_units * kRN, "hipMalloc (unit sums)")) ||
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
          static_cast<uint32_t>(st->ncb), sb.keys[0], st->maxbits + 1);
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
      unit_assign_kernel<<<1, 1024, 0, s>>>(st->units, st->unit_count, kUnitGroups,
                                            st->wg_first);
      PM_CHECK_LAUNCH("unit_assign_kernel");
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
                  0, s, st->units, st->wg_first, idx, px, py, pz, mass, n, dx, inv_vol,        \
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
                    st->eflag, st->esc_any, st->dirty, st->parity, full, st->node_count,
                    st->maxbits, bl, n, st->nbn, rho);
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
