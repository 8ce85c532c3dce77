// Stable LSD radix sort of 32-bit keys carrying 32-bit particle indices (pm_deposit.hip),
// shared by the sorted deposit (cell keys) and the brick binning of the bricked deposit.
#pragma once

#include "pm_common.h"

namespace pm {
namespace radix {

constexpr int kTileKeys = 4096;  // keys per workgroup tile
constexpr int kMaxDigitBits = 9; // widest digit of one pass

inline int64_t tiles_of(int64_t np) { return (np + kTileKeys - 1) / kTileKeys; }
// Entries of the per-tile digit histogram for up to np keys.
inline int64_t hist_words(int64_t np) { return tiles_of(np) << kMaxDigitBits; }

// Digit widths for keys below `limit` (limit itself included): the fewest passes of at
// most kMaxDigitBits bits, widths as even as possible. Returns the number of passes.
int plan_digits(uint64_t limit, int bits[4]);

// Sorts keys_in[0..np) stably by the digits bits[0..passes) taken from bit `shift` upward,
// carrying the input positions (the first pass sorts the identity; no index array is
// read). keys[2], vals[2]: ping-pong buffers of np entries; hist: hist_words(np) words;
// digit: 2^kMaxDigitBits words. *out is the index of the buffers holding the result.
// Each pass records the deposit.radix_hist / radix_scan / radix_scatter timers.
pm_status sort(const uint32_t* keys_in, uint32_t* const keys[2], int32_t* const vals[2],
               uint32_t* hist, uint32_t* digit, int64_t np, int shift, const int* bits,
               int passes, int* out, hipStream_t s);

}  // namespace radix
}  // namespace pm
