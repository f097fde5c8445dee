// Command-line driver with the reference's contract (src/main.cu:66-257, run.sh:8):
//   ac_main data.bin pattern.bin > result
// prints one "%e" line per MU value, data-vertex-major (dn x pn, src/main.cu:247,257), and "elapsed=<us>" on stderr
// covering prep + loop + output (src/main.cu:104-105,249-250).  The lines are formatted in parallel, in chunks written in
// order, each value by the same snprintf("%e\n"), so the bytes are those of one printf per value.  Extensions, none of
// which changes the default output:
//   --csv  dataV.csv dataE.csv patternV.csv patternE.csv   read the CSV quartet instead of the two .bin files
//   --iters N                                                iteration count (default: pattern vertices)
//   --device D                                               HIP device
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ac_mi355x.h"

static int die(const char* what, const char* msg) {
  std::fprintf(stderr, "ac_main: %s: %s\n", what, msg ? msg : "");
  return 1;
}

// Writes one "%e\n" line per value.  Up to 16 threads each format a chunk of 2^20 values into their own buffer; the
// buffers of one round of chunks are written in chunk order before the next round starts.
static void write_values(const float* v, size_t total, FILE* out) {
  // A float's "%e\n" has a two-digit exponent: at most 14 characters, 15 bytes with snprintf's terminator.
  constexpr size_t kChunk = size_t(1) << 20, kLine = 16;
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t nchunks = (total + kChunk - 1) / kChunk;
  std::vector<std::vector<char>> buf(nt, std::vector<char>(std::min(total, kChunk) * kLine));
  std::vector<size_t> used(nt);
  for (size_t c0 = 0; c0 < nchunks; c0 += nt) {
    const size_t cn = std::min<size_t>(nt, nchunks - c0);
    auto format = [&](size_t k) {
      const size_t b = (c0 + k) * kChunk, e = std::min(total, b + kChunk);
      char* p = buf[k].data();
      size_t n = 0;
      for (size_t i = b; i < e; ++i) n += (size_t)std::snprintf(p + n, kLine, "%e\n", (double)v[i]);
      used[k] = n;
    };
    std::vector<std::thread> th;
    for (size_t k = 1; k < cn; ++k) th.emplace_back(format, k);
    format(0);
    for (auto& t : th) t.join();
    for (size_t k = 0; k < cn; ++k) std::fwrite(buf[k].data(), 1, used[k], out);
  }
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
  write_values(mu.data(), total, stdout);
  ac_destroy(ctx);
  ac_graph_free(&data);
  ac_graph_free(&patt);
  return 0;
}
