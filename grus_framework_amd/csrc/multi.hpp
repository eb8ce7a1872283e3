// multi.hpp -- batched multi-source traversal engines (up to 64 lanes).
//
// No reference counterpart: the reference runs one source per process
// (samples/bfs/bfs.cu:135-173, --src flag).  Here up to 64 queries share one
// traversal, one bit per query ("lane") in a u64 per vertex (MS-BFS, Then et
// al., VLDB 2014).  An engine steps all busy lanes together; a lane whose
// frontier empties is finished, its slot can be reset and given to a new
// query between steps while the other lanes keep running (continuous
// batching).  Every lane computes exactly its single-source result:
//   BFS  level[lane][v] = (step of discovery) + 1 - (step the lane started)
//   SSSP dist[v][lane]  = Bellman-Ford fixed point of its own source
// independent of which other lanes share the batch.
// SSSP rows start 16 bits wide (64 lanes = one 128-byte line per vertex,
// 0xFFFF = unreached).  A relaxation that would store a distance of 0xFFFF
// or more into an unreached entry flags its lane; the engine then widens
// every row to 32 bits and restarts the flagged lanes from their sources
// (a Bellman-Ford fixed point does not depend on when a lane starts), so
// every result is exact either way.
#pragma once

#include <vector>

#include "graph.hpp"

namespace mgg {

enum class MsKind { BFS = 0, SSSP = 1 };

// digest of a device row of V u32 (see grus_result_digest); synchronous
void row_digest(Context &ctx, const uint32_t *row, uint64_t V, uint64_t *reached,
                uint64_t *checksum);

struct MsCounters {
  uint64_t alive;          // lanes whose new frontier is not empty this step
  uint64_t front_edges;   // out-edges of the new frontier (direction heuristic)
  uint32_t qcount;        // frontier queue length
  uint32_t long_count;    // pull: vertices handed to wave-cooperative scan
  uint64_t overflow;      // SSSP, 16-bit rows: lanes with a distance past 0xFFFE
  uint64_t pad[3];
};

class MsEngine {
 public:
  MsEngine(Graph &g, MsKind kind, int lanes, bool own_stream);
  ~MsEngine();
  MsKind kind() const { return kind_; }
  int lanes() const { return lanes_; }
  int free_lanes() const;
  uint64_t busy_mask() const { return busy_; }
  bool active() const { return busy_ != 0; }
  // admit sources into free lanes (between steps); returns the lanes used
  std::vector<int> admit(const uint32_t *srcs, int n);
  // one expansion step of every busy lane; returns lanes that finished
  std::vector<int> step();
  // lane result (V u32) into a device row, then optionally to host memory;
  // reached = finite entries, checksum = sum over v of mix64(v << 32 | value)
  void extract(int lane, uint32_t *host_out, uint64_t *reached, uint64_t *checksum,
               uint32_t *device_out = nullptr);
  // reached / checksum (as extract) of every lane in mask, in one pass over
  // the engine's state; results land at [lane]
  void digest(uint64_t mask, uint64_t *reached, uint64_t *checksum);
  void release(int lane);  // lane becomes free (reset lazily at next admit)
  Context &ctx() { return *ctx_; }
  uint64_t steps() const { return t_; }
  uint64_t state_bytes() const;
  bool narrow() const { return narrow_; }  // SSSP rows of 16 bits

 private:
  Graph &g_;
  MsKind kind_;
  int lanes_;
  Context *ctx_;
  std::unique_ptr<Context> own_ctx_;
  uint64_t busy_ = 0;      // lanes holding a query
  uint64_t finished_mask_ = 0;  // busy lanes whose frontier has emptied
  uint64_t dirty_ = 0;     // lanes whose bits must be cleared before reuse
  uint32_t start_[64] = {};
  uint32_t src_[64] = {};  // each lane's source (a widened SSSP lane restarts from it)
  bool narrow_ = false;
  uint64_t t_ = 0;         // global step counter
  uint64_t qn_ = 0, front_edges_ = 0;
  Buf visit_, front_, next_, vals_, queue_, longq_, row_, digest_, srcs_dev_, cnt_, push_cnt_;
  Buf hmirror_;  // pinned host words: the counters as published, then read-back scratch
  PushWork push_;
  void reset_lanes(uint64_t mask);
  void inject(const uint32_t *srcs, const int *lanes, int n);
  // 16-bit SSSP rows to 32 bits; the lanes in `restart` start over
  void widen(uint64_t restart);
  MsCounters publish();
  uint64_t read_u64(const uint64_t *d);
};

}  // namespace mgg
