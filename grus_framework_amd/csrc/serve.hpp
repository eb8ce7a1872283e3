// serve.hpp -- continuous-batching query server (no reference counterpart).
#pragma once

#include "graph.hpp"

namespace mgg {

struct Server {
  virtual ~Server() = default;
  virtual uint64_t submit(int kind, uint32_t src, uint32_t *out_host) = 0;
  virtual void poll(grus_query_result *res, int timeout_ms) = 0;
  virtual void stats(grus_server_stats *out) = 0;
};

std::unique_ptr<Server> make_server(Graph &g, const grus_server_opts *opts);

}  // namespace mgg
