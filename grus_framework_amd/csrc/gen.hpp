// gen.hpp -- synthetic graph specification shared by the device generators.
//
// The reference has no generator (it reads .gr files only), so the RMAT and
// grid inputs are this project's own specification:
//   RMAT sample i, level l (0 = most significant bit): a 32-bit uniform r
//   taken from h_j = mix64(key + ((i << 5) | j) * GOLDEN), j = l / 2, low
//   half for even l, high half for odd l, key = mix64(seed ^ K_RMAT).
//   Quadrant r < t0 -> (0,0), r < t1 -> (0,1), r < t2 -> (1,0), else
//   (1,1); the first bit extends u, the second v.  t0 = a*2^32,
//   t1 = (a+b)*2^32, t2 = (a+b+c)*2^32, truncated (computed in double on
//   the host).  Self-loops are dropped, every pair is added in both
//   directions, duplicates removed, rows sorted by destination.
//   Weight of the undirected pair {u,v}: wlo + mix64(wkey + (min<<32|max))
//   mod (whi-wlo+1), wkey = mix64(wseed ^ K_WEIGHT), so both directions
//   carry the same weight.
//   Grid rows x cols: id = r*cols + c, neighbours in ascending id order
//   (up, left, right, down).
#pragma once

#include "device.hpp"

namespace mgg {

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;
constexpr uint64_t K_RMAT = 0x6A09E667F3BCC909ull;
constexpr uint64_t K_WEIGHT = 0xBB67AE8584CAA73Bull;

struct RmatSpec {
  int scale;
  uint32_t t0, t1, t2;
  uint64_t key;
};

inline RmatSpec rmat_spec(int scale, double a, double b, double c, uint64_t seed) {
  RmatSpec s;
  s.scale = scale;
  s.t0 = (uint32_t)(a * 4294967296.0);
  s.t1 = (uint32_t)((a + b) * 4294967296.0);
  s.t2 = (uint32_t)((a + b + c) * 4294967296.0);
  s.key = mix64(seed ^ K_RMAT);
  return s;
}

__host__ __device__ __forceinline__ void rmat_sample(const RmatSpec &s, uint64_t i, uint32_t &u,
                                                     uint32_t &v) {
  uint32_t uu = 0, vv = 0;
  uint64_t h = 0;
  for (int l = 0; l < s.scale; l++) {
    if ((l & 1) == 0) h = mix64(s.key + (((i << 5) | (uint64_t)(l >> 1)) * GOLDEN));
    uint32_t r = (l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
    uint32_t q = (r < s.t0) ? 0u : (r < s.t1) ? 1u : (r < s.t2) ? 2u : 3u;
    uu = (uu << 1) | (q >> 1);
    vv = (vv << 1) | (q & 1u);
  }
  u = uu;
  v = vv;
}

__host__ __device__ __forceinline__ uint64_t weight_key(uint64_t wseed) {
  return mix64(wseed ^ K_WEIGHT);
}

__host__ __device__ __forceinline__ uint32_t edge_weight(uint32_t u, uint32_t v, uint64_t wkey,
                                                         uint32_t wlo, uint32_t whi) {
  uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
  uint64_t h = mix64(wkey + ((lo << 32) | hi));
  return wlo + (uint32_t)(h % (uint64_t)(whi - wlo + 1));
}

}  // namespace mgg
