// prims.hpp -- device-wide primitives (rocPRIM instantiated once, in prims.hip).
#pragma once

#include "graph.hpp"

namespace mgg {

// out[i] = sum of in[0..i), for i in [0, n); in and out must not alias.
void scan_excl_u64(Context &ctx, const uint64_t *in, uint64_t *out, size_t n);
// The same with caller-provided temporary storage (no allocation): raises
// GRUS_E_BUDGET when tmp_bytes is below scan_excl_u64_bytes(n).
void scan_excl_u64_with(Context &ctx, const uint64_t *in, uint64_t *out, size_t n, void *tmp,
                        size_t tmp_bytes);
size_t scan_excl_u64_bytes(Context &ctx, size_t n);
// out[i] = sum of in[0..i) in 32 bits (the caller guarantees the total fits).
void scan_excl_u32(Context &ctx, const uint32_t *in, uint32_t *out, size_t n);
// Exclusive scan of the 0/1 "first of its run and not the sentinel" flags of
// sorted keys: pos[j] = number of distinct non-sentinel keys before j.
void scan_unique_flags(Context &ctx, const uint64_t *keys, uint64_t *pos, size_t n,
                       uint64_t sentinel);
// Sort keys ascending using a and b as double buffer; returns the buffer
// that holds the result.
uint64_t *sort_keys_u64(Context &ctx, uint64_t *a, uint64_t *b, size_t n, int end_bit = 64);
// Sort (key, value) pairs; returns key result buffer, *vals_out the values.
uint64_t *sort_pairs_u64_u32(Context &ctx, uint64_t *ka, uint64_t *kb, uint32_t *va,
                             uint32_t *vb, size_t n, uint32_t **vals_out, int end_bit = 64);

}  // namespace mgg
