// pending.cpp -- entry points whose engines are not built yet.
#include "apps.hpp"
#include "serve.hpp"

namespace mgg {
void run_bfs_multi(Graph &, const uint32_t *, int, uint32_t *, grus_stats *) {
  raise(GRUS_E_UNSUPPORTED, "multi-source bfs not built");
}
void run_sssp_multi(Graph &, const uint32_t *, int, uint32_t *, grus_stats *) {
  raise(GRUS_E_UNSUPPORTED, "multi-source sssp not built");
}
std::unique_ptr<Server> make_server(Graph &, const grus_server_opts *) {
  raise(GRUS_E_UNSUPPORTED, "server not built");
}
}  // namespace mgg
