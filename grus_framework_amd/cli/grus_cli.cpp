// grus_cli.cpp -- reference-compatible command line for the four apps.
//
// Mirrors samples/{bfs,sssp,pagerank,cc}/main.cu + Skeleton<App>
// (reference include/app.cuh:33-43): the same gflags-style flag names and
// meanings (include/app.cuh:11-31), the "itr N in T" timing line
// (samples/bfs/bfs.cu:170) and the --output dump "<vertex> <value>" per line
// (print::SaveResults, include/print.cuh:36-45).  It is a host program over
// the C-ABI of libgrus_hip.so (include/grus_c.h); no engine code here.
//
// The app is chosen by the executable name (grus_bfs, grus_sssp, grus_pr,
// grus_cc) or by --app.  Inputs: --input <file.gr> (Galois .gr, as the
// reference), or a synthetic --rmat=<scale> / --grid=<rows>x<cols>.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/grus_c.h"

namespace {

struct Flags {
  std::map<std::string, std::string> kv;
  bool has(const std::string &k) const { return kv.count(k) != 0; }
  std::string str(const std::string &k, const std::string &d) const {
    auto it = kv.find(k);
    return it == kv.end() ? d : it->second;
  }
  long long num(const std::string &k, long long d) const {
    return has(k) ? std::atoll(kv.at(k).c_str()) : d;
  }
  double dbl(const std::string &k, double d) const {
    return has(k) ? std::atof(kv.at(k).c_str()) : d;
  }
  bool flag(const std::string &k, bool d) const {
    if (!has(k)) return d;
    const std::string &v = kv.at(k);
    return v.empty() || v == "1" || v == "true" || v == "yes";
  }
};

// gflags syntax: --name=value, --name value, --name (true), --noname (false)
const char *kBoolFlags[] = {"pull", "pf", "ab", "rm", "pl", "opt", "v", "symmetric", "weighted",
                            "pr_unnormalized"};

bool is_bool(const std::string &n) {
  for (const char *b : kBoolFlags)
    if (n == b) return true;
  return false;
}

Flags parse(int argc, char **argv) {
  Flags f;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0 && a.rfind("-", 0) == 0) a = "-" + a;
    if (a.rfind("--", 0) != 0) {
      fprintf(stderr, "unexpected argument: %s\n", argv[i]);
      exit(2);
    }
    a = a.substr(2);
    size_t eq = a.find('=');
    if (eq != std::string::npos) {
      f.kv[a.substr(0, eq)] = a.substr(eq + 1);
    } else if (is_bool(a)) {
      f.kv[a] = "true";
    } else if (a.rfind("no", 0) == 0 && is_bool(a.substr(2))) {
      f.kv[a.substr(2)] = "false";
    } else if (i + 1 < argc) {
      f.kv[a] = argv[++i];
    } else {
      f.kv[a] = "";
    }
  }
  return f;
}

void die(int rc, const char *what) {
  fprintf(stderr, "%s failed: %s (status %d)\n", what, grus_last_error(), rc);
  exit(1);
}
#define CHECK(call)                   \
  do {                                \
    int rc_ = (call);                 \
    if (rc_ != GRUS_OK) die(rc_, #call); \
  } while (0)

void usage(const char *prog) {
  fprintf(stderr,
          "usage: %s [--app bfs|bfs_2s|sssp|pr|cc] (--input g.gr | --rmat S | --grid RxC)\n"
          "  reference flags: --output FILE --src N --device N --ngpu 1 --pull\n"
          "                   --opt --pf --ab --rm --pl --wl_th X (accepted, unused)\n"
          "  placement:       --placement hbm|managed|capped --hbm_cap_gb G\n"
          "                   --mem_limit_gb G (capped: whole-device cap)\n"
          "  co-location:     --cu_limit N (run on the first N compute units)\n"
          "  inputs:          --symmetric --edge_factor N --seed N --weights LO,HI --wseed N\n"
          "  apps:            --direction auto|push|pull (bfs)  --iters N --alpha A (pr)\n"
          "                   --stage_edges N (bfs_2s: staging buffer, edges)\n"
          "                   --method hook|lp (cc)  --pr_unnormalized (dump ranks x V)\n"
          "  output:          --precise_output (ranks with 9 digits instead of 6)\n",
          prog);
}

}  // namespace

int main(int argc, char **argv) {
  Flags f = parse(argc, argv);
  std::string prog = argv[0];
  std::string base = prog.substr(prog.find_last_of('/') + 1);
  std::string app = f.str("app", base.rfind("grus_", 0) == 0 ? base.substr(5) : "bfs");
  if (app == "pagerank") app = "pr";
  if (f.has("help") ||
      (app != "bfs" && app != "bfs_2s" && app != "sssp" && app != "pr" && app != "cc")) {
    usage(argv[0]);
    return f.has("help") ? 0 : 2;
  }
  if (f.num("ngpu", 1) != 1) {
    fprintf(stderr, "--ngpu: this build runs on one GPU\n");
    return 2;
  }

  // placement: --opt (default on, include/app.cuh:31) keeps everything in HBM
  // when it fits, else places what fits (capped at the free device memory);
  // --pf/--ab/--rm/--pl without --opt select managed memory with those hints
  grus_place place{};
  std::string pm = f.str("placement", "");
  bool hints = f.flag("pf", false) || f.flag("ab", false) || f.flag("rm", false) ||
               f.flag("pl", false);
  if (pm.empty()) pm = (!f.flag("opt", true) && hints) ? "managed" : "hbm";
  if (pm == "managed") place.mode = GRUS_PLACE_MANAGED;
  else if (pm == "capped") place.mode = GRUS_PLACE_CAPPED;
  else place.mode = GRUS_PLACE_HBM;
  place.hbm_cap_bytes = (uint64_t)(f.dbl("hbm_cap_gb", 0.0) * (1ull << 30));
  // --mem_limit_gb: the whole device footprint, as the reference's
  // tools/mem_limit.cu caps the device before a run (capped placement)
  place.device_budget_bytes = (uint64_t)(f.dbl("mem_limit_gb", 0.0) * (1ull << 30));
  place.prefetch = f.flag("pf", false);
  place.accessed_by = f.flag("ab", false);
  place.read_mostly = f.flag("rm", false);
  place.preferred_location = f.flag("pl", false);

  grus_ctx *ctx = nullptr;
  CHECK(grus_init((int)f.num("device", 0), &ctx));
  if (f.has("cu_limit")) {  // the reference's SM-limit experiments, as a CU mask
    int active = 0, total = 0;
    CHECK(grus_cu_count(ctx, &active, &total));
    const long n = f.num("cu_limit", total);
    if (n < 1) {
      fprintf(stderr, "--cu_limit: at least one compute unit\n");
      return 2;
    }
    if (n < total) {
      std::vector<uint32_t> mask((total + 31) / 32, 0u);
      for (long i = 0; i < n; i++) mask[i / 32] |= 1u << (i % 32);
      CHECK(grus_set_cu_mask(ctx, mask.data(), (uint32_t)mask.size()));
    }
  }
  grus_graph *g = nullptr;
  bool need_w = (app == "sssp");
  uint32_t wlo = 0, whi = 0;
  if (f.has("weights")) sscanf(f.str("weights", "").c_str(), "%u,%u", &wlo, &whi);
  if (f.has("input")) {
    CHECK(grus_graph_load_gr(ctx, f.str("input", "").c_str(), need_w, f.flag("symmetric", false),
                             &place, &g));
  } else if (f.has("rmat")) {
    if (need_w && !wlo) wlo = 1, whi = 255;
    CHECK(grus_graph_gen_rmat(ctx, (int)f.num("rmat", 16), (int)f.num("edge_factor", 16), 0.57,
                              0.19, 0.19, (uint64_t)f.num("seed", 1), wlo, whi,
                              (uint64_t)f.num("wseed", 3), &place, &g));
  } else if (f.has("grid")) {
    unsigned r = 0, c = 0;
    sscanf(f.str("grid", "").c_str(), "%ux%u", &r, &c);
    if (need_w && !wlo) wlo = 1, whi = 1000;
    CHECK(grus_graph_gen_grid(ctx, r, c, wlo, whi, (uint64_t)f.num("wseed", 2), &place, &g));
  } else {
    usage(argv[0]);
    return 2;
  }
  grus_graph_info info;
  CHECK(grus_graph_info_get(g, &info));
  // the reference's loader report (include/graph_loader.cuh:81-82, 135-137)
  printf("graph has %llu nodes and %llu edges\n", (unsigned long long)info.V,
         (unsigned long long)info.E);
  printf("%u has max out degree %llu\n", info.max_degree_vertex,
         (unsigned long long)info.max_degree);

  std::vector<uint32_t> u32;
  std::vector<float> f32;
  grus_stats st{};
  uint32_t src = (uint32_t)f.num("src", 0);
  auto t0 = std::chrono::steady_clock::now();
  if (app == "bfs") {
    std::string d = f.str("direction", f.flag("pull", false) ? "pull" : "auto");
    int dir = d == "push" ? GRUS_DIR_PUSH : d == "pull" ? GRUS_DIR_PULL : GRUS_DIR_AUTO;
    u32.resize(info.V);
    CHECK(grus_bfs(g, src, dir, u32.data(), &st));
  } else if (app == "bfs_2s") {  // samples/bfs_2s: two-stage active-subgraph streaming
    u32.resize(info.V);
    CHECK(grus_bfs_2s(g, src, (uint64_t)f.num("stage_edges", 0), u32.data(), &st));
  } else if (app == "sssp") {
    u32.resize(info.V);
    CHECK(grus_sssp(g, src, u32.data(), &st));
  } else if (app == "cc") {
    u32.resize(info.V);
    CHECK(grus_cc(g, f.str("method", "hook") == "lp" ? 1 : 0, u32.data(), &st));
  } else {
    f32.resize(info.V);
    CHECK(grus_pagerank(g, (float)f.dbl("alpha", 0.85), (int)f.num("iters", 20), f32.data(), &st));
  }
  double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
  // "itr N in T" as samples/*/*.cu print it, T = device milliseconds
  printf("itr %u in %.3f\n", st.iterations, st.device_ms);
  printf("{\"app\": \"%s\", \"V\": %llu, \"E\": %llu, \"iterations\": %u, \"device_ms\": %.4f, "
         "\"wall_ms\": %.4f, \"vertices_reached\": %llu, \"edges_reached\": %llu, "
         "\"gteps\": %.3f}\n",
         app.c_str(), (unsigned long long)info.V, (unsigned long long)info.E, st.iterations,
         st.device_ms, wall, (unsigned long long)st.vertices_reached,
         (unsigned long long)st.edges_reached,
         st.device_ms > 0 ? (double)st.edges_reached / st.device_ms / 1e6 : 0.0);
  if (f.has("output")) {
    std::string path = f.str("output", "");
    FILE *out = fopen(path.c_str(), "w");
    if (!out) {
      fprintf(stderr, "cannot open %s\n", path.c_str());
      return 1;
    }
    printf("Saving the results into the following file:\n>> %s\n", path.c_str());
    double scale = f.flag("pr_unnormalized", false) ? (double)info.V : 1.0;
    // ranks as the reference's SaveResults streams them (an ofstream at its
    // default precision: %g, 6 significant digits, include/print.cuh:36-45);
    // --precise_output keeps 9 digits, enough to round-trip an FP32 value
    const char *ffmt = f.flag("precise_output", false) ? "%llu %.9g\n" : "%llu %g\n";
    for (uint64_t i = 0; i < info.V; i++) {
      if (!u32.empty()) fprintf(out, "%llu %u\n", (unsigned long long)i, u32[i]);
      else fprintf(out, ffmt, (unsigned long long)i, (double)(float)(f32[i] * scale));
    }
    fclose(out);
    printf("Done saving.\n");
  }
  grus_graph_destroy(g);
  grus_fini(ctx);
  return 0;
}
