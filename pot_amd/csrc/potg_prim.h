// potg_prim.h -- device primitives (scan, compaction, radix sort) used by
// the step pipeline.  All launches go to the given stream; none allocates.
#pragma once

#include "potg_dev.h"

namespace potg {

// words of scratch the scan needs for n elements
size_t scan_temp_words(uint32_t n);
// out[i] = add + sum(in[0..i)); n_dev (optional) caps the count on the
// device; total/total2 (optional, device) receive the sum.  in may equal out.
void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, const uint32_t* n_dev,
                    uint32_t* total, uint32_t* total2, uint32_t* temp, uint32_t add,
                    hipStream_t s);

size_t radix_temp_words(uint32_t n);
// Stable LSD sort of (keys, vals) on bits [0, end_bit); the sorted data end
// up in keys/vals (the pointers are swapped with the _alt buffers as needed).
// ipt: elements per thread of a tile, 8 or 16 (tiles of 2048 or 4096).
void radix_sort_pairs(uint64_t*& keys, uint64_t*& keys_alt, uint32_t*& vals, uint32_t*& vals_alt,
                      uint32_t n, int end_bit, uint32_t* temp, int ipt, hipStream_t s);

void iota(uint32_t* v, uint32_t n, hipStream_t s);
// out[k] = i for the k-th i with flags[i] != 0, in order; *count = total
void compact_flags(const uint32_t* flags, uint32_t* pos, uint32_t n, uint32_t* out,
                   uint32_t* count, uint32_t* temp, hipStream_t s);

}  // namespace potg
