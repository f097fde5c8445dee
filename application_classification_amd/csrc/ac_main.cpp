// Command-line driver with the reference's contract (src/main.cu:66-257, run.sh:8):
//   ac_main data.bin pattern.bin > result
// prints one "%e" line per MU value, data-vertex-major (dn x pn, src/main.cu:247,257), and "elapsed=<us>" on stderr
// covering prep + loop + output (src/main.cu:104-105,249-250).  The graphs are read at the reference's own precision
// (FP64 features, uint64 ids) and handed to the library as they are (ac_set_graph64).  The lines are formatted in
// parallel, in chunks written in order, each value by the same snprintf("%e\n"), so the bytes are those of one printf
// per value.  Extensions, none of which changes the default output:
//   --csv  dataV.csv dataE.csv patternV.csv patternE.csv   read the CSV quartet instead of the two .bin files
//   --to-bin vertex.csv edges.csv out.bin                   convert a CSV pair to the .bin prob2bin.py writes
//                                                            (prob2bin.py:11-64, byte for byte), then exit
//   --iters N                                                iteration count (default: pattern vertices)
//   --device D                                               HIP device
//   --mu-bin FILE                                            also write MU in binary: uint64 dn, uint64 pn, then
//                                                            dn x pn float32 in the printed order
//   --topk K FILE                                            also write, per pattern vertex p, its K most likely data
//                                                            vertices as lines "p rank v mu" (ties: lower v first)
//   --no-text                                                skip the "%e" lines (with --mu-bin / --topk)
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

static int to_bin(const char* vcsv, const char* ecsv, const char* out) {
  ac_graph64 g;
  if (ac_graph64_load_csv(vcsv, ecsv, &g)) return die("--to-bin", ac_global_error());
  const int rc = ac_graph64_save_bin(out, &g);
  ac_graph64_free(&g);
  if (rc) return die("--to-bin", ac_global_error());
  return 0;
}

static int write_mu_bin(const char* path, const float* mu, uint64_t dn, uint64_t pn) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return die("--mu-bin", path);
  const uint64_t hdr[2] = {dn, pn};
  bool ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1;
  ok = ok && (dn * pn == 0 || std::fwrite(mu, sizeof(float), dn * pn, f) == dn * pn);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : die("--mu-bin", "write failed");
}

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  int iters = -1, device = 0, topk = 0;
  bool csv = false, text = true;
  const char* mu_bin = nullptr;
  const char* topk_file = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--csv")) csv = true;
    else if (!std::strcmp(argv[i], "--to-bin") && i + 3 < argc) return to_bin(argv[i + 1], argv[i + 2], argv[i + 3]);
    else if (!std::strcmp(argv[i], "--iters") && i + 1 < argc) iters = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) device = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--mu-bin") && i + 1 < argc) mu_bin = argv[++i];
    else if (!std::strcmp(argv[i], "--topk") && i + 2 < argc) {
      topk = std::atoi(argv[++i]);
      topk_file = argv[++i];
    } else if (!std::strcmp(argv[i], "--no-text")) text = false;
    else pos.push_back(argv[i]);
  }
  if ((!csv && pos.size() != 2) || (csv && pos.size() != 4) || topk < 0) {
    std::fprintf(stderr,
                 "usage: ac_main data.bin pattern.bin\n"
                 "       ac_main --csv dataV.csv dataE.csv patternV.csv patternE.csv\n"
                 "       ac_main --to-bin vertex.csv edges.csv out.bin\n"
                 "options: --iters N  --device D  --mu-bin FILE  --topk K FILE  --no-text\n");
    return 2;
  }
  ac_graph64 data, patt;
  int rc = csv ? ac_graph64_load_csv(pos[0].c_str(), pos[1].c_str(), &data) : ac_graph64_load_bin(pos[0].c_str(), &data);
  if (rc) return die("data graph", ac_global_error());
  rc = csv ? ac_graph64_load_csv(pos[2].c_str(), pos[3].c_str(), &patt) : ac_graph64_load_bin(pos[1].c_str(), &patt);
  if (rc) return die("pattern graph", ac_global_error());
  ac_ctx* ctx = nullptr;
  if (ac_create(&ctx, device)) return die("ac_create", ac_global_error());
  if (ac_set_graph64(ctx, AC_DATA, &data) || ac_set_graph64(ctx, AC_PATTERN, &patt))
    return die("ac_set_graph64", ac_last_error(ctx));
  if (ac_run(ctx, iters)) return die("ac_run", ac_last_error(ctx));
  double prep_ms = 0, loop_ms = 0;
  ac_get_timing(ctx, &prep_ms, &loop_ms);
  std::fprintf(stderr, "elapsed=%lld\n", (long long)(1000.0 * (prep_ms + loop_ms)));
  const uint64_t dn = data.num_nodes, pn = patt.num_nodes;
  std::vector<float> mu((size_t)(dn * pn));
  if (ac_get_mu(ctx, mu.data())) return die("ac_get_mu", ac_last_error(ctx));
  if (topk > 0) {
    const int k = (int)std::min<uint64_t>((uint64_t)topk, dn);
    std::vector<int32_t> idx((size_t)pn * k);
    std::vector<float> val((size_t)pn * k);
    if (ac_get_topk(ctx, k, idx.data(), val.data())) return die("ac_get_topk", ac_last_error(ctx));
    FILE* f = std::fopen(topk_file, "w");
    if (!f) return die("--topk", topk_file);
    for (uint64_t p = 0; p < pn; ++p)
      for (int r = 0; r < k; ++r) std::fprintf(f, "%llu %d %d %e\n", (unsigned long long)p, r, idx[p * k + r], (double)val[p * k + r]);
    if (std::fclose(f) != 0) return die("--topk", "write failed");
  }
  if (mu_bin && write_mu_bin(mu_bin, mu.data(), dn, pn)) return 1;
  if (text) write_values(mu.data(), mu.size(), stdout);
  ac_destroy(ctx);
  ac_graph64_free(&data);
  ac_graph64_free(&patt);
  return 0;
}
