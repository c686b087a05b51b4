// Data-parallel primitives used by the simplifier: exclusive scan and a stable LSD radix sort.
// Sizes come either from the host (n) or from a device word (n_dev != nullptr, with n as the
// upper bound that sizes the grid), so a pass can be enqueued without a round trip.
#pragma once

#include "common.h"

namespace mhip
{

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems; // 4096

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems; // 4096

// temporary space (in uint32 words) needed by scan_exclusive for n elements
size_t scan_temp_words(size_t n);

// out[i] = sum(in[0..i)), total (optional) = sum(in[0..n)); out may alias in
void scan_exclusive(hipStream_t s, const uint32_t* in, uint32_t* out, size_t n, const uint32_t* n_dev, uint32_t* total, uint32_t* temp);

// temporary space (in uint32 words) needed by radix_sort_pairs for n elements
size_t sort_temp_words(size_t n);

// stable sort of (keys, vals) by the low `bits` bits of the key; the sorted pairs end in
// (keys, vals) on return (the alt buffers are ping-pong space of the same size)
void radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n, const uint32_t* n_dev, int bits, uint32_t* temp);

void fill_u32(hipStream_t s, uint32_t* p, uint32_t value, size_t n);
void iota_u32(hipStream_t s, uint32_t* p, size_t n);

} // namespace mhip
