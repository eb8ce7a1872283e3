// multi.hpp -- batched multi-source traversal engines.
//
// No reference counterpart: the reference runs one source per process
// (samples/bfs/bfs.cu:135-173, --src flag).  Here many queries share one
// traversal as bits of a u64 mask per vertex (MS-BFS, Then et al., VLDB
// 2014).  A query occupies a slot; an engine steps all busy slots together;
// a slot whose frontier empties is finished and can be given to a new query
// between steps while the others keep running (continuous batching).  Every
// slot computes exactly its single-source result:
//   BFS  level[slot][v] = (step of discovery) + 1 - (step the slot started)
//   SSSP dist[v][slot]  = Bellman-Ford fixed point of its own source
// independent of which other slots share the batch.
//
// BFS: 64 slots, one per mask bit ("lane").
// SSSP: 128 slots in 64 lanes of two.  A vertex's distances are one row
// of 128 entries, lane p owning slots 2p and 2p+1: 16-bit entries (a 32-bit
// word per lane, 0xFFFF = unreached; a 256-byte row) until some slot needs
// a distance of 0xFFFF or more, then 32-bit entries (a 64-bit pair per
// lane).  A relaxation that would store such a distance into an unreached
// entry flags its slot; the engine widens every row and restarts the
// flagged slots from their sources (a Bellman-Ford fixed point does not
// depend on when a slot starts), so every result is exact either way.  The
// frontier masks are per lane: a lane is active at v when either of its
// slots improved there; relaxing the other slot again changes nothing.
#pragma once

#include <vector>

#include "graph.hpp"

namespace mgg {

enum class MsKind { BFS = 0, SSSP = 1 };

// digest of a device row of V u32 (see grus_result_digest); synchronous
void row_digest(Context &ctx, const uint32_t *row, uint64_t V, uint64_t *reached,
                uint64_t *checksum);

struct MsCounters {
  uint64_t alive;          // BFS: lanes whose new frontier is not empty this step
  uint64_t front_edges;   // out-edges of the new frontier (direction heuristic)
  uint32_t qcount;        // frontier queue length
  uint32_t long_count;    // pull: vertices handed to wave-cooperative scan
  uint64_t slot_alive[2];  // SSSP: slots 2p / 2p+1 (bit p) that improved a distance
  uint64_t overflow[2];    // SSSP, 16-bit rows: slots 2p / 2p+1 that need more bits
  uint64_t pad;
};

// a set of query slots (up to 128)
using SlotMask = unsigned __int128;
inline SlotMask slot_bit(int s) { return (SlotMask)1 << s; }
inline int slot_count(SlotMask m) {
  return __builtin_popcountll((uint64_t)m) + __builtin_popcountll((uint64_t)(m >> 64));
}
inline int slot_lowest(SlotMask m) {
  const uint64_t lo = (uint64_t)m;
  return lo ? __builtin_ctzll(lo) : 64 + __builtin_ctzll((uint64_t)(m >> 64));
}

class MsEngine {
 public:
  // lanes: 1..64 (BFS: one slot each; SSSP: two slots each).  own_stream:
  // a context (stream) of its own, else the graph's
  MsEngine(Graph &g, MsKind kind, int lanes, bool own_stream);
  // the same on a context the caller owns (it must outlive the engine)
  MsEngine(Graph &g, MsKind kind, int lanes, Context *ctx);
  ~MsEngine();
  MsKind kind() const { return kind_; }
  int slots() const { return slots_; }
  int free_slots() const { return slots_ - slot_count(busy_); }
  SlotMask busy_mask() const { return busy_; }
  bool active() const { return busy_ != 0; }
  // admit sources into free slots (between steps); returns the slots used
  std::vector<int> admit(const uint32_t *srcs, int n);
  // one expansion step of every busy slot; returns slots that finished
  std::vector<int> step();
  // slot result (V u32) into a device row, then optionally to host memory;
  // reached = finite entries, checksum = sum over v of mix64(v << 32 | value)
  void extract(int slot, uint32_t *host_out, uint64_t *reached, uint64_t *checksum,
               uint32_t *device_out = nullptr);
  // reached / checksum (as extract) of every slot in mask, in one pass over
  // the engine's state; results land at [slot]
  void digest(SlotMask mask, uint64_t *reached, uint64_t *checksum);
  void release(int slot);  // slot becomes free (reset lazily at next admit)
  Context &ctx() { return *ctx_; }
  uint64_t steps() const { return t_; }
  uint64_t state_bytes() const;
  bool narrow() const { return narrow_; }  // SSSP rows of 16-bit entries

 private:
  Graph &g_;
  MsKind kind_;
  int lanes_, slots_;
  Context *ctx_;
  std::unique_ptr<Context> own_ctx_;
  SlotMask busy_ = 0;      // slots holding a query
  SlotMask finished_ = 0;  // busy slots whose frontier has emptied
  SlotMask dirty_ = 0;     // slots whose state must be reset before reuse
  uint32_t start_[128] = {};
  uint32_t src_[128] = {};  // each slot's source (a widened SSSP slot restarts from it)
  bool narrow_ = false;
  uint64_t t_ = 0;         // global step counter
  uint64_t qn_ = 0, front_edges_ = 0;
  Buf visit_, front_, next_, vals_, queue_, longq_, row_, digest_, srcs_dev_, cnt_, push_cnt_;
  Buf hmirror_;  // pinned host words: the counters as published, then read-back scratch
  PushWork push_;
  Buf scan_tmp_;  // the step's scan temporaries, owned: a step allocates nothing
  SlotMask all_slots() const {
    return slots_ == 128 ? ~(SlotMask)0 : (slot_bit(slots_) - 1);
  }
  uint64_t lanes_of(SlotMask m) const;  // lanes holding a slot of m
  void reset_slots(SlotMask mask);
  void inject(const uint32_t *srcs, const int *slots, int n);
  // 16-bit SSSP rows to 32 bits; the slots in `restart` start over
  void widen(SlotMask restart);
  MsCounters publish();
  uint64_t read_u64(const uint64_t *d);
  void init();  // device state, on ctx_
};

}  // namespace mgg
