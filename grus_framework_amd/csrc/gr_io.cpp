// gr_io.cpp -- Galois .gr v1 reader/writer (host only).
//
// Byte layout follows graph_loader::ReadGraphGR (reference
// include/graph_loader.cuh:65-159): u64 header {version=1, sizeEdgeTy, N,
// M} (:71-80), N u64 end offsets (:114-120), M u32 destinations
// (:122-131), a 4-byte pad when M is odd and the file is weighted
// (:139-141), then M u32 weights (:142-150).  Unlike the reference, offsets
// stay u64 and off[0] is written, short reads are errors rather than
// messages, and an unweighted file yields weights of 1 only when asked
// (:95-101).
#include <cstdio>
#include <cstring>
#include <vector>

#include "graph.hpp"

namespace mgg {

namespace {
struct File {
  FILE *f = nullptr;
  File(const std::string &p, const char *mode) : f(fopen(p.c_str(), mode)) {
    if (!f) raise(GRUS_E_IO, "cannot open " + p);
  }
  ~File() {
    if (f) fclose(f);
  }
  void read(void *dst, size_t sz, size_t n, const char *what) {
    // chunked so that multi-GB arrays do not hit short-read limits
    char *d = static_cast<char *>(dst);
    size_t left = n * sz;
    while (left) {
      size_t chunk = left < (size_t(1) << 30) ? left : (size_t(1) << 30);
      size_t got = fread(d, 1, chunk, f);
      if (got != chunk) raise(GRUS_E_FORMAT, std::string("truncated .gr file: ") + what);
      d += chunk;
      left -= chunk;
    }
  }
  void write(const void *src, size_t sz, size_t n) {
    if (n && fwrite(src, sz, n, f) != n) raise(GRUS_E_IO, "short write to .gr file");
  }
};
}  // namespace

GrHeader gr_read_header(const std::string &path) {
  File f(path, "rb");
  uint64_t h[4];
  f.read(h, sizeof(uint64_t), 4, "header");
  if (h[0] != 1) raise(GRUS_E_FORMAT, "unknown .gr version " + std::to_string(h[0]));
  if (h[1] != 0 && h[1] != 4)
    raise(GRUS_E_FORMAT, "unsupported .gr edge data size " + std::to_string(h[1]));
  if (h[2] > 0xFFFFFFFFull) raise(GRUS_E_FORMAT, ".gr vertex count exceeds u32 ids");
  return GrHeader{h[2], h[3], h[1]};
}

void gr_read(const std::string &path, uint64_t *off, uint32_t *col, uint32_t *w) {
  GrHeader h = gr_read_header(path);
  File f(path, "rb");
  if (fseek(f.f, 32, SEEK_SET) != 0) raise(GRUS_E_FORMAT, "seek failed");
  off[0] = 0;
  f.read(off + 1, sizeof(uint64_t), h.V, "offsets");
  f.read(col, sizeof(uint32_t), h.E, "destinations");
  if (w) {
    if (h.size_edge_ty) {
      if (h.E % 2) {
        uint32_t pad;
        f.read(&pad, 4, 1, "pad");
      }
      f.read(w, sizeof(uint32_t), h.E, "weights");
    } else {
      for (uint64_t i = 0; i < h.E; i++) w[i] = 1;
    }
  }
  if (off[h.V] != h.E) raise(GRUS_E_FORMAT, ".gr offsets do not end at the edge count");
  for (uint64_t i = 0; i < h.V; i++)
    if (off[i] > off[i + 1]) raise(GRUS_E_FORMAT, ".gr offsets are not monotone");
  for (uint64_t i = 0; i < h.E; i++)
    if (col[i] >= h.V) raise(GRUS_E_FORMAT, ".gr destination out of range");
}

void gr_write(const std::string &path, uint64_t V, uint64_t E, const uint64_t *off,
              const uint32_t *col, const uint32_t *w) {
  if (off[0] != 0 || off[V] != E) raise(GRUS_E_INVALID, "offsets must run from 0 to E");
  File f(path, "wb");
  uint64_t h[4] = {1, w ? 4u : 0u, V, E};
  f.write(h, sizeof(uint64_t), 4);
  f.write(off + 1, sizeof(uint64_t), V);
  f.write(col, sizeof(uint32_t), E);
  if (w) {
    if (E % 2) {
      uint32_t pad = 0;
      f.write(&pad, 4, 1);
    }
    f.write(w, sizeof(uint32_t), E);
  }
}

}  // namespace mgg
