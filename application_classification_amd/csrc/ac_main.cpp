// Command-line driver with the reference's contract (src/main.cu:66-257, run.sh:8):
//   ac_main data.bin pattern.bin > result
// prints one "%e" line per MU value, data-vertex-major (dn x pn, src/main.cu:247,257), and "elapsed=<us>" on stderr
// covering prep + loop + output (src/main.cu:104-105,249-250).  Extensions, none of which changes the default output:
//   --csv  dataV.csv dataE.csv patternV.csv patternE.csv   read the CSV quartet instead of the two .bin files
//   --iters N                                                iteration count (default: pattern vertices)
//   --device D                                               HIP device
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ac_mi355x.h"

static int die(const char* what, const char* msg) {
  std::fprintf(stderr, "ac_main: %s: %s\n", what, msg ? msg : "");
  return 1;
}

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  int iters = -1, device = 0;
  bool csv = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--csv")) csv = true;
    else if (!std::strcmp(argv[i], "--iters") && i + 1 < argc) iters = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) device = std::atoi(argv[++i]);
    else pos.push_back(argv[i]);
  }
  if ((!csv && pos.size() != 2) || (csv && pos.size() != 4)) {
    std::fprintf(stderr, "usage: ac_main data.bin pattern.bin\n       ac_main --csv dataV.csv dataE.csv patternV.csv patternE.csv\n");
    return 2;
  }
  ac_host_graph data, patt;
  int rc = csv ? ac_graph_load_csv(pos[0].c_str(), pos[1].c_str(), &data) : ac_graph_load_bin(pos[0].c_str(), &data);
  if (rc) return die("data graph", ac_global_error());
  rc = csv ? ac_graph_load_csv(pos[2].c_str(), pos[3].c_str(), &patt) : ac_graph_load_bin(pos[1].c_str(), &patt);
  if (rc) return die("pattern graph", ac_global_error());
  ac_ctx* ctx = nullptr;
  if (ac_create(&ctx, device)) return die("ac_create", ac_global_error());
  if (ac_set_graph(ctx, AC_DATA, data.num_nodes, data.node_feat_dim, data.node_feats, data.num_edges,
                   data.edge_feat_dim, data.edge_feats, data.srcs, data.dsts) ||
      ac_set_graph(ctx, AC_PATTERN, patt.num_nodes, patt.node_feat_dim, patt.node_feats, patt.num_edges,
                   patt.edge_feat_dim, patt.edge_feats, patt.srcs, patt.dsts))
    return die("ac_set_graph", ac_last_error(ctx));
  if (ac_run(ctx, iters)) return die("ac_run", ac_last_error(ctx));
  double prep_ms = 0, loop_ms = 0;
  ac_get_timing(ctx, &prep_ms, &loop_ms);
  std::fprintf(stderr, "elapsed=%lld\n", (long long)(1000.0 * (prep_ms + loop_ms)));
  const size_t total = (size_t)data.num_nodes * (size_t)patt.num_nodes;
  std::vector<float> mu(total);
  if (ac_get_mu(ctx, mu.data())) return die("ac_get_mu", ac_last_error(ctx));
  std::vector<char> buf(1 << 20);
  size_t used = 0;
  for (size_t i = 0; i < total; ++i) {
    if (used + 64 > buf.size()) {
      std::fwrite(buf.data(), 1, used, stdout);
      used = 0;
    }
    used += (size_t)std::snprintf(buf.data() + used, 64, "%e\n", (double)mu[i]);
  }
  std::fwrite(buf.data(), 1, used, stdout);
  ac_destroy(ctx);
  ac_graph_free(&data);
  ac_graph_free(&patt);
  return 0;
}
