// prims.hip -- rocPRIM device-wide scan and radix sort behind small wrappers.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "prims.hpp"

namespace mgg {

namespace {
struct UniqueFlag {
  const uint64_t *k;
  uint64_t sentinel;
  __host__ __device__ uint64_t operator()(size_t j) const {
    uint64_t x = k[j];
    return (x != sentinel && (j == 0 || k[j - 1] != x)) ? 1ull : 0ull;
  }
};

void *tmp(Context &ctx, size_t bytes) {
  ctx.scan_tmp.ensure(bytes < 256 ? 256 : bytes);
  return ctx.scan_tmp.p;
}
}  // namespace

void scan_excl_u64(Context &ctx, const uint64_t *in, uint64_t *out, size_t n) {
  if (n == 0) return;
  size_t bytes = 0;
  GRUS_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, uint64_t(0), n,
                                   rocprim::plus<uint64_t>(), ctx.stream));
  void *t = tmp(ctx, bytes);
  GRUS_HIP(rocprim::exclusive_scan(t, bytes, in, out, uint64_t(0), n, rocprim::plus<uint64_t>(),
                                   ctx.stream));
}

size_t scan_excl_u64_bytes(Context &ctx, size_t n) {
  size_t bytes = 0;
  GRUS_HIP(rocprim::exclusive_scan(nullptr, bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                   uint64_t(0), n, rocprim::plus<uint64_t>(), ctx.stream));
  return bytes;
}

void scan_excl_u64_with(Context &ctx, const uint64_t *in, uint64_t *out, size_t n, void *tmp,
                        size_t tmp_bytes) {
  if (n == 0) return;
  size_t bytes = 0;
  GRUS_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, uint64_t(0), n,
                                   rocprim::plus<uint64_t>(), ctx.stream));
  if (bytes > tmp_bytes) raise(GRUS_E_BUDGET, "scan temporaries exceed the given storage");
  GRUS_HIP(rocprim::exclusive_scan(tmp, bytes, in, out, uint64_t(0), n, rocprim::plus<uint64_t>(),
                                   ctx.stream));
}

void scan_excl_u32(Context &ctx, const uint32_t *in, uint32_t *out, size_t n) {
  if (n == 0) return;
  size_t bytes = 0;
  GRUS_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, uint32_t(0), n,
                                   rocprim::plus<uint32_t>(), ctx.stream));
  void *t = tmp(ctx, bytes);
  GRUS_HIP(rocprim::exclusive_scan(t, bytes, in, out, uint32_t(0), n, rocprim::plus<uint32_t>(),
                                   ctx.stream));
}

void scan_unique_flags(Context &ctx, const uint64_t *keys, uint64_t *pos, size_t n,
                       uint64_t sentinel) {
  if (n == 0) return;
  auto it = rocprim::make_transform_iterator(rocprim::make_counting_iterator<size_t>(0),
                                             UniqueFlag{keys, sentinel});
  size_t bytes = 0;
  GRUS_HIP(rocprim::exclusive_scan(nullptr, bytes, it, pos, uint64_t(0), n,
                                   rocprim::plus<uint64_t>(), ctx.stream));
  void *t = tmp(ctx, bytes);
  GRUS_HIP(rocprim::exclusive_scan(t, bytes, it, pos, uint64_t(0), n, rocprim::plus<uint64_t>(),
                                   ctx.stream));
}

uint64_t *sort_keys_u64(Context &ctx, uint64_t *a, uint64_t *b, size_t n, int end_bit) {
  if (n == 0) return a;
  rocprim::double_buffer<uint64_t> db(a, b);
  size_t bytes = 0;
  GRUS_HIP(rocprim::radix_sort_keys(nullptr, bytes, db, n, 0, end_bit, ctx.stream));
  void *t = tmp(ctx, bytes);
  GRUS_HIP(rocprim::radix_sort_keys(t, bytes, db, n, 0, end_bit, ctx.stream));
  return db.current();
}

uint64_t *sort_pairs_u64_u32(Context &ctx, uint64_t *ka, uint64_t *kb, uint32_t *va,
                             uint32_t *vb, size_t n, uint32_t **vals_out, int end_bit) {
  if (n == 0) {
    *vals_out = va;
    return ka;
  }
  rocprim::double_buffer<uint64_t> dk(ka, kb);
  rocprim::double_buffer<uint32_t> dv(va, vb);
  size_t bytes = 0;
  GRUS_HIP(rocprim::radix_sort_pairs(nullptr, bytes, dk, dv, n, 0, end_bit, ctx.stream));
  void *t = tmp(ctx, bytes);
  GRUS_HIP(rocprim::radix_sort_pairs(t, bytes, dk, dv, n, 0, end_bit, ctx.stream));
  *vals_out = dv.current();
  return dk.current();
}

}  // namespace mgg
