// File formats of the reference's driver: the binary graph layout (prob2bin.py:46-64, read by
// src/main.cu:21-63) and the space-separated CSV pair it is converted from (prob2bin.py:11-39).
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ac_mi355x.h"
#include "ac_host.hpp"

namespace ac {
namespace {
thread_local std::string g_error;
}
void set_global_error(const std::string& msg) { g_error = msg; }
const char* global_error() { return g_error.c_str(); }
}  // namespace ac

using namespace ac;

extern "C" const char* ac_global_error(void) { return ac::global_error(); }

extern "C" void ac_graph_free(ac_host_graph* g) {
  if (!g) return;
  std::free(g->node_feats);
  std::free(g->edge_feats);
  std::free(g->srcs);
  std::free(g->dsts);
  std::memset(g, 0, sizeof(*g));
}

namespace {

bool alloc_graph(ac_host_graph* g, int64_t nV, int64_t dV, int64_t nE, int64_t dE) {
  std::memset(g, 0, sizeof(*g));
  g->num_nodes = nV;
  g->node_feat_dim = dV;
  g->num_edges = nE;
  g->edge_feat_dim = dE;
  g->node_feats = (float*)std::malloc(sizeof(float) * std::max<int64_t>(1, nV * dV));
  g->edge_feats = (float*)std::malloc(sizeof(float) * std::max<int64_t>(1, nE * dE));
  g->srcs = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(1, nE));
  g->dsts = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(1, nE));
  if (!g->node_feats || !g->edge_feats || !g->srcs || !g->dsts) {
    ac_graph_free(g);
    return false;
  }
  return true;
}

int check_ids(const ac_host_graph* g, const char* where) {
  for (int64_t e = 0; e < g->num_edges; ++e)
    if (g->srcs[e] < 0 || g->srcs[e] >= g->num_nodes || g->dsts[e] < 0 || g->dsts[e] >= g->num_nodes) {
      set_global_error(std::string(where) + ": edge " + std::to_string(e) + " has an endpoint outside [0, " +
                       std::to_string(g->num_nodes) + ")");
      return AC_ERR_IO;
    }
  return AC_OK;
}

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// Reads n doubles from the file and narrows them to float.
bool read_f64_as_f32(FILE* f, float* out, int64_t n) {
  const int64_t chunk = 1 << 20;
  std::vector<double> buf((size_t)std::min<int64_t>(std::max<int64_t>(n, 1), chunk));
  for (int64_t i = 0; i < n; i += chunk) {
    int64_t k = std::min(chunk, n - i);
    if ((int64_t)std::fread(buf.data(), sizeof(double), k, f) != k) return false;
    for (int64_t j = 0; j < k; ++j) out[i + j] = (float)buf[j];
  }
  return true;
}

bool read_u64_as_i32(FILE* f, int32_t* out, int64_t n, int64_t limit) {
  const int64_t chunk = 1 << 20;
  std::vector<uint64_t> buf((size_t)std::min<int64_t>(std::max<int64_t>(n, 1), chunk));
  for (int64_t i = 0; i < n; i += chunk) {
    int64_t k = std::min(chunk, n - i);
    if ((int64_t)std::fread(buf.data(), sizeof(uint64_t), k, f) != k) return false;
    for (int64_t j = 0; j < k; ++j) {
      if (buf[j] >= (uint64_t)limit) return false;
      out[i + j] = (int32_t)buf[j];
    }
  }
  return true;
}

}  // namespace

extern "C" int ac_graph_load_bin(const char* path, ac_host_graph* out) {
  if (!path || !out) {
    set_global_error("ac_graph_load_bin: null argument");
    return AC_ERR_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  File file(path, "rb");
  if (!file.f) {
    set_global_error(std::string("ac_graph_load_bin: cannot open ") + path + ": " + std::strerror(errno));
    return AC_ERR_IO;
  }
  uint64_t hdr[4];
  if (std::fread(hdr, sizeof(uint64_t), 4, file.f) != 4) {
    set_global_error(std::string("ac_graph_load_bin: truncated header in ") + path);
    return AC_ERR_IO;
  }
  const uint64_t lim = (uint64_t)INT32_MAX;
  if (hdr[0] == 0 || hdr[0] > lim || hdr[1] > 4096 || hdr[2] > ((uint64_t)1 << 40) || hdr[3] > 4096) {
    set_global_error(std::string("ac_graph_load_bin: implausible header in ") + path);
    return AC_ERR_IO;
  }
  // The file length must be exactly what the header announces.
  std::fseek(file.f, 0, SEEK_END);
  long long len = std::ftell(file.f);
  std::fseek(file.f, 32, SEEK_SET);
  const long long want =
      32 + 8LL * (long long)(hdr[0] * hdr[1] + hdr[2] * hdr[3] + 2 * hdr[2]);
  if (len != want) {
    set_global_error(std::string("ac_graph_load_bin: ") + path + " has " + std::to_string(len) +
                     " bytes, header implies " + std::to_string(want));
    return AC_ERR_IO;
  }
  if (!alloc_graph(out, (int64_t)hdr[0], (int64_t)hdr[1], (int64_t)hdr[2], (int64_t)hdr[3])) {
    set_global_error("ac_graph_load_bin: out of host memory");
    return AC_ERR_OOM;
  }
  bool ok = read_f64_as_f32(file.f, out->node_feats, out->num_nodes * out->node_feat_dim) &&
            read_f64_as_f32(file.f, out->edge_feats, out->num_edges * out->edge_feat_dim) &&
            read_u64_as_i32(file.f, out->srcs, out->num_edges, out->num_nodes) &&
            read_u64_as_i32(file.f, out->dsts, out->num_edges, out->num_nodes);
  if (!ok) {
    ac_graph_free(out);
    set_global_error(std::string("ac_graph_load_bin: short read or edge id >= num_nodes in ") + path);
    return AC_ERR_IO;
  }
  return AC_OK;
}

extern "C" int ac_graph_save_bin(const char* path, const ac_host_graph* g) {
  if (!path || !g) {
    set_global_error("ac_graph_save_bin: null argument");
    return AC_ERR_ARG;
  }
  File file(path, "wb");
  if (!file.f) {
    set_global_error(std::string("ac_graph_save_bin: cannot open ") + path);
    return AC_ERR_IO;
  }
  uint64_t hdr[4] = {(uint64_t)g->num_nodes, (uint64_t)g->node_feat_dim, (uint64_t)g->num_edges,
                     (uint64_t)g->edge_feat_dim};
  bool ok = std::fwrite(hdr, sizeof(hdr), 1, file.f) == 1;
  auto put_f = [&](const float* x, int64_t n) {
    std::vector<double> buf;
    for (int64_t i = 0; i < n && ok; i += 1 << 20) {
      int64_t k = std::min<int64_t>(1 << 20, n - i);
      buf.assign(x + i, x + i + k);
      ok = (int64_t)std::fwrite(buf.data(), sizeof(double), k, file.f) == k;
    }
  };
  auto put_i = [&](const int32_t* x, int64_t n) {
    std::vector<uint64_t> buf;
    for (int64_t i = 0; i < n && ok; i += 1 << 20) {
      int64_t k = std::min<int64_t>(1 << 20, n - i);
      buf.assign(x + i, x + i + k);
      ok = (int64_t)std::fwrite(buf.data(), sizeof(uint64_t), k, file.f) == k;
    }
  };
  put_f(g->node_feats, g->num_nodes * g->node_feat_dim);
  put_f(g->edge_feats, g->num_edges * g->edge_feat_dim);
  put_i(g->srcs, g->num_edges);
  put_i(g->dsts, g->num_edges);
  if (!ok) {
    set_global_error(std::string("ac_graph_save_bin: write failed for ") + path);
    return AC_ERR_IO;
  }
  return AC_OK;
}

namespace {

// Powers of ten 10^0 .. 10^308 as correctly rounded doubles.
const double* pow10_table() {
  static const std::vector<double> t = [] {
    std::vector<double> v(309);
    char buf[16];
    for (int k = 0; k <= 308; ++k) {
      std::snprintf(buf, sizeof(buf), "1e%d", k);
      v[k] = std::strtod(buf, nullptr);
    }
    return v;
  }();
  return t.data();
}

// Decimal text to double as pandas.read_csv converts it by default (prob2bin.py:12 reads the CSVs with it; pandas
// 2.3.3 here, the reference pins no version), so that converted .bin files are byte-identical to prob2bin.py's:
// up to 17 significant digits are accumulated in a double (later integer digits only raise the exponent, later
// fraction digits are dropped), the sign is applied, and the value is scaled once by a table power of ten
// (dividing twice below 1e-308).  This is not always the correctly rounded value: about one value in ten of the
// reference's georgiy CSVs differs from strtod's by one ulp.  Returns the end of the token in *end (== p: none).
double parse_decimal(const char* p, const char** end) {
  const char* s = p;
  bool neg = false;
  if (*s == '-' || *s == '+') neg = *s++ == '-';
  constexpr int kMaxDigits = 17;
  double num = 0.0;
  int exp10 = 0, ndig = 0;
  const char* d0 = s;
  for (; *s >= '0' && *s <= '9'; ++s) {
    if (ndig < kMaxDigits) {
      num = num * 10.0 + (double)(*s - '0');
      ++ndig;
    } else {
      ++exp10;
    }
  }
  bool any = s != d0;
  if (*s == '.') {
    ++s;
    int nfrac = 0;
    for (; *s >= '0' && *s <= '9'; ++s) {
      any = true;
      if (ndig < kMaxDigits) {
        num = num * 10.0 + (double)(*s - '0');
        ++ndig;
        ++nfrac;
      }
    }
    exp10 -= nfrac;
  }
  if (!any) {
    *end = p;
    return NAN;
  }
  if (neg) num = -num;
  if (*s == 'e' || *s == 'E') {
    const char* e = s + 1;
    bool eneg = false;
    if (*e == '-' || *e == '+') eneg = *e++ == '-';
    int n = 0, k = 0;
    for (; k < kMaxDigits && *e >= '0' && *e <= '9'; ++e, ++k) n = n * 10 + (*e - '0');
    if (k > 0) {
      s = e;
      exp10 += eneg ? -n : n;
    }
  }
  *end = s;
  const double* P = pow10_table();
  if (exp10 > 308) return neg ? -HUGE_VAL : HUGE_VAL;
  if (exp10 > 0) return num * P[exp10];
  if (exp10 < -308) return exp10 < -616 ? 0.0 : num / P[-308 - exp10] / P[308];
  return num / P[-exp10];
}

// Whole-file reader with a cursor; tokens are separated by spaces, tabs, CR or LF.
struct Text {
  std::string buf;
  size_t pos = 0;
  bool load(const char* path) {
    File f(path, "rb");
    if (!f.f) return false;
    std::fseek(f.f, 0, SEEK_END);
    long len = std::ftell(f.f);
    std::fseek(f.f, 0, SEEK_SET);
    buf.resize(len > 0 ? (size_t)len : 0);
    return len <= 0 || std::fread(&buf[0], 1, buf.size(), f.f) == buf.size();
  }
  void skip_line() {
    while (pos < buf.size() && buf[pos] != '\n') ++pos;
    if (pos < buf.size()) ++pos;
  }
  void skip_blank_lines() {
    for (;;) {
      size_t p = pos;
      while (p < buf.size() && (buf[p] == ' ' || buf[p] == '\t' || buf[p] == '\r')) ++p;
      if (p < buf.size() && buf[p] == '\n') {
        pos = p + 1;
        continue;
      }
      return;
    }
  }
  bool at_end() {
    skip_blank_lines();
    return pos >= buf.size();
  }
  // Parses one row into `vals`; returns false at end of file.
  bool row(std::vector<double>& vals) {
    vals.clear();
    skip_blank_lines();
    if (pos >= buf.size()) return false;
    while (pos < buf.size() && buf[pos] != '\n') {
      while (pos < buf.size() && (buf[pos] == ' ' || buf[pos] == '\t' || buf[pos] == '\r')) ++pos;
      if (pos >= buf.size() || buf[pos] == '\n') break;
      const char* end = nullptr;
      const double v = parse_decimal(buf.c_str() + pos, &end);
      const size_t stop = (size_t)(end - buf.c_str());
      const bool whole = stop < buf.size() ? (buf[stop] == ' ' || buf[stop] == '\n' || buf[stop] == '\t' ||
                                                buf[stop] == '\r')
                                             : true;
      if (stop == pos || !whole) {  // not a number: NaN, then skip the token
        vals.push_back(NAN);
        while (pos < buf.size() && buf[pos] != ' ' && buf[pos] != '\n' && buf[pos] != '\t' && buf[pos] != '\r') ++pos;
      } else {
        vals.push_back(v);
        pos = stop;
      }
    }
    if (pos < buf.size()) ++pos;
    return true;
  }
};

}  // namespace

extern "C" int ac_graph_load_csv(const char* vertex_csv, const char* edge_csv, ac_host_graph* out) {
  if (!vertex_csv || !edge_csv || !out) {
    set_global_error("ac_graph_load_csv: null argument");
    return AC_ERR_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  Text vt, et;
  if (!vt.load(vertex_csv) || !et.load(edge_csv)) {
    set_global_error(std::string("ac_graph_load_csv: cannot read ") + vertex_csv + " / " + edge_csv);
    return AC_ERR_IO;
  }
  vt.skip_line();  // header "N C" (prob2bin.py:12 skips it)
  et.skip_line();
  std::vector<std::vector<double>> vrows, erows;
  std::vector<double> r;
  while (vt.row(r)) vrows.push_back(r);
  while (et.row(r)) erows.push_back(r);
  if (vrows.empty() || vrows[0].size() < 1) {
    set_global_error(std::string("ac_graph_load_csv: no vertex rows in ") + vertex_csv);
    return AC_ERR_IO;
  }
  const int64_t nV = (int64_t)vrows.size(), dV = (int64_t)vrows[0].size() - 1;
  const int64_t nE = (int64_t)erows.size(), dE = nE ? (int64_t)erows[0].size() - 2 : 0;
  if (nE && dE < 0) {
    set_global_error(std::string("ac_graph_load_csv: edge rows need src and dst in ") + edge_csv);
    return AC_ERR_IO;
  }
  if (!alloc_graph(out, nV, dV, nE, dE)) {
    set_global_error("ac_graph_load_csv: out of host memory");
    return AC_ERR_OOM;
  }
  for (int64_t i = 0; i < nV; ++i) {
    // The id column must equal the row index (prob2bin.py:28-29).
    if ((int64_t)vrows[i].size() != dV + 1 || vrows[i][0] != (double)i) {
      ac_graph_free(out);
      set_global_error(std::string("ac_graph_load_csv: vertex row ") + std::to_string(i) +
                       " is ragged or its id is not the row index");
      return AC_ERR_IO;
    }
    for (int64_t k = 0; k < dV; ++k) out->node_feats[i * dV + k] = (float)vrows[i][k + 1];
  }
  for (int64_t e = 0; e < nE; ++e) {
    const auto& row = erows[e];
    if ((int64_t)row.size() != dE + 2 || row[0] != std::floor(row[0]) || row[1] != std::floor(row[1])) {
      ac_graph_free(out);
      set_global_error(std::string("ac_graph_load_csv: edge row ") + std::to_string(e) + " is malformed");
      return AC_ERR_IO;
    }
    out->srcs[e] = row[0] >= 0 && row[0] < nV ? (int32_t)row[0] : -1;
    out->dsts[e] = row[1] >= 0 && row[1] < nV ? (int32_t)row[1] : -1;
    for (int64_t k = 0; k < dE; ++k) out->edge_feats[e * dE + k] = (float)row[k + 2];
  }
  int rc = check_ids(out, "ac_graph_load_csv");
  if (rc != AC_OK) ac_graph_free(out);
  return rc;
}

extern "C" int ac_graph_save_csv(const char* vertex_csv, const char* edge_csv, const ac_host_graph* g) {
  if (!vertex_csv || !edge_csv || !g) {
    set_global_error("ac_graph_save_csv: null argument");
    return AC_ERR_ARG;
  }
  File vf(vertex_csv, "wb"), ef(edge_csv, "wb");
  if (!vf.f || !ef.f) {
    set_global_error("ac_graph_save_csv: cannot open output files");
    return AC_ERR_IO;
  }
  std::fprintf(vf.f, "%lld %lld\n", (long long)g->num_nodes, (long long)g->node_feat_dim + 1);
  for (int64_t i = 0; i < g->num_nodes; ++i) {
    std::fprintf(vf.f, "%lld", (long long)i);
    for (int64_t k = 0; k < g->node_feat_dim; ++k)
      std::fprintf(vf.f, " %.17g", (double)g->node_feats[i * g->node_feat_dim + k]);
    std::fputc('\n', vf.f);
  }
  std::fprintf(ef.f, "%lld %lld\n", (long long)g->num_edges, (long long)g->edge_feat_dim + 2);
  for (int64_t e = 0; e < g->num_edges; ++e) {
    std::fprintf(ef.f, "%d %d", g->srcs[e], g->dsts[e]);
    for (int64_t k = 0; k < g->edge_feat_dim; ++k)
      std::fprintf(ef.f, " %.17g", (double)g->edge_feats[e * g->edge_feat_dim + k]);
    std::fputc('\n', ef.f);
  }
  return AC_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// The reference's own host graph at full precision (src/main.cu:8-19: uint64 sizes and ids, FP64 features).

extern "C" void ac_graph64_free(ac_graph64* g) {
  if (!g) return;
  std::free(g->node_feats);
  std::free(g->edge_feats);
  std::free(g->srcs);
  std::free(g->dsts);
  std::memset(g, 0, sizeof(*g));
}

namespace {

bool alloc_graph64(ac_graph64* g, uint64_t nV, uint64_t dV, uint64_t nE, uint64_t dE) {
  std::memset(g, 0, sizeof(*g));
  g->num_nodes = nV;
  g->node_feat_dim = dV;
  g->num_edges = nE;
  g->edge_feat_dim = dE;
  g->node_feats = (double*)std::malloc(sizeof(double) * std::max<uint64_t>(1, nV * dV));
  g->edge_feats = (double*)std::malloc(sizeof(double) * std::max<uint64_t>(1, nE * dE));
  g->srcs = (uint64_t*)std::malloc(sizeof(uint64_t) * std::max<uint64_t>(1, nE));
  g->dsts = (uint64_t*)std::malloc(sizeof(uint64_t) * std::max<uint64_t>(1, nE));
  if (!g->node_feats || !g->edge_feats || !g->srcs || !g->dsts) {
    ac_graph64_free(g);
    return false;
  }
  return true;
}

// Reads the header of a .bin (four uint64: nV, dV, nE, dE) and checks that the file length is exactly what it
// announces; the cursor is left on the first feature.
int read_bin_header(FILE* f, const char* path, const char* who, uint64_t hdr[4]) {
  if (std::fread(hdr, sizeof(uint64_t), 4, f) != 4) {
    set_global_error(std::string(who) + ": truncated header in " + path);
    return AC_ERR_IO;
  }
  if (hdr[0] == 0 || hdr[0] > (uint64_t)INT32_MAX || hdr[1] > 4096 || hdr[2] > ((uint64_t)1 << 40) || hdr[3] > 4096) {
    set_global_error(std::string(who) + ": implausible header in " + path);
    return AC_ERR_IO;
  }
  std::fseek(f, 0, SEEK_END);
  const long long len = std::ftell(f);
  std::fseek(f, 32, SEEK_SET);
  const long long want = 32 + 8LL * (long long)(hdr[0] * hdr[1] + hdr[2] * hdr[3] + 2 * hdr[2]);
  if (len != want) {
    set_global_error(std::string(who) + ": " + path + " has " + std::to_string(len) + " bytes, header implies " +
                     std::to_string(want));
    return AC_ERR_IO;
  }
  return AC_OK;
}

template <typename T>
bool read_all(FILE* f, T* out, uint64_t n) {
  return n == 0 || std::fread(out, sizeof(T), n, f) == n;
}

template <typename T>
bool write_all(FILE* f, const T* in, uint64_t n) {
  return n == 0 || std::fwrite(in, sizeof(T), n, f) == n;
}

}  // namespace

extern "C" int ac_graph64_load_bin(const char* path, ac_graph64* out) {
  if (!path || !out) {
    set_global_error("ac_graph64_load_bin: null argument");
    return AC_ERR_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  File file(path, "rb");
  if (!file.f) {
    set_global_error(std::string("ac_graph64_load_bin: cannot open ") + path + ": " + std::strerror(errno));
    return AC_ERR_IO;
  }
  uint64_t hdr[4];
  int rc = read_bin_header(file.f, path, "ac_graph64_load_bin", hdr);
  if (rc != AC_OK) return rc;
  if (!alloc_graph64(out, hdr[0], hdr[1], hdr[2], hdr[3])) {
    set_global_error("ac_graph64_load_bin: out of host memory");
    return AC_ERR_OOM;
  }
  const bool ok = read_all(file.f, out->node_feats, hdr[0] * hdr[1]) && read_all(file.f, out->edge_feats, hdr[2] * hdr[3]) &&
                  read_all(file.f, out->srcs, hdr[2]) && read_all(file.f, out->dsts, hdr[2]);
  if (!ok) {
    ac_graph64_free(out);
    set_global_error(std::string("ac_graph64_load_bin: short read in ") + path);
    return AC_ERR_IO;
  }
  for (uint64_t e = 0; e < hdr[2]; ++e)
    if (out->srcs[e] >= hdr[0] || out->dsts[e] >= hdr[0]) {
      set_global_error(std::string("ac_graph64_load_bin: edge ") + std::to_string(e) + " has an endpoint >= " +
                       std::to_string(hdr[0]) + " in " + path);
      ac_graph64_free(out);
      return AC_ERR_IO;
    }
  return AC_OK;
}

extern "C" int ac_graph64_save_bin(const char* path, const ac_graph64* g) {
  if (!path || !g) {
    set_global_error("ac_graph64_save_bin: null argument");
    return AC_ERR_ARG;
  }
  File file(path, "wb");
  if (!file.f) {
    set_global_error(std::string("ac_graph64_save_bin: cannot open ") + path);
    return AC_ERR_IO;
  }
  const uint64_t hdr[4] = {g->num_nodes, g->node_feat_dim, g->num_edges, g->edge_feat_dim};
  const bool ok = write_all(file.f, hdr, 4) && write_all(file.f, g->node_feats, g->num_nodes * g->node_feat_dim) &&
                  write_all(file.f, g->edge_feats, g->num_edges * g->edge_feat_dim) &&
                  write_all(file.f, g->srcs, g->num_edges) && write_all(file.f, g->dsts, g->num_edges);
  if (!ok) {
    set_global_error(std::string("ac_graph64_save_bin: write failed for ") + path);
    return AC_ERR_IO;
  }
  return AC_OK;
}

extern "C" int ac_graph64_load_csv(const char* vertex_csv, const char* edge_csv, ac_graph64* out) {
  if (!vertex_csv || !edge_csv || !out) {
    set_global_error("ac_graph64_load_csv: null argument");
    return AC_ERR_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  Text vt, et;
  if (!vt.load(vertex_csv) || !et.load(edge_csv)) {
    set_global_error(std::string("ac_graph64_load_csv: cannot read ") + vertex_csv + " / " + edge_csv);
    return AC_ERR_IO;
  }
  vt.skip_line();  // header "N C" (prob2bin.py:12 skips it)
  et.skip_line();
  std::vector<std::vector<double>> vrows, erows;
  std::vector<double> r;
  while (vt.row(r)) vrows.push_back(r);
  while (et.row(r)) erows.push_back(r);
  if (vrows.empty() || vrows[0].empty()) {
    set_global_error(std::string("ac_graph64_load_csv: no vertex rows in ") + vertex_csv);
    return AC_ERR_IO;
  }
  const uint64_t nV = vrows.size(), dV = vrows[0].size() - 1;
  const uint64_t nE = erows.size();
  if (nE && erows[0].size() < 2) {
    set_global_error(std::string("ac_graph64_load_csv: edge rows need src and dst in ") + edge_csv);
    return AC_ERR_IO;
  }
  const uint64_t dE = nE ? erows[0].size() - 2 : 0;
  if (!alloc_graph64(out, nV, dV, nE, dE)) {
    set_global_error("ac_graph64_load_csv: out of host memory");
    return AC_ERR_OOM;
  }
  for (uint64_t i = 0; i < nV; ++i) {
    // The id column must equal the row index (prob2bin.py:28-29).
    if (vrows[i].size() != dV + 1 || vrows[i][0] != (double)i) {
      ac_graph64_free(out);
      set_global_error(std::string("ac_graph64_load_csv: vertex row ") + std::to_string(i) +
                       " is ragged or its id is not the row index");
      return AC_ERR_IO;
    }
    std::copy(vrows[i].begin() + 1, vrows[i].end(), out->node_feats + i * dV);
  }
  for (uint64_t e = 0; e < nE; ++e) {
    const auto& row = erows[e];
    if (row.size() != dE + 2 || !(row[0] >= 0 && row[0] < (double)nV && row[0] == std::floor(row[0])) ||
        !(row[1] >= 0 && row[1] < (double)nV && row[1] == std::floor(row[1]))) {
      ac_graph64_free(out);
      set_global_error(std::string("ac_graph64_load_csv: edge row ") + std::to_string(e) +
                       " is ragged or has an endpoint outside the vertex rows");
      return AC_ERR_IO;
    }
    out->srcs[e] = (uint64_t)row[0];
    out->dsts[e] = (uint64_t)row[1];
    std::copy(row.begin() + 2, row.end(), out->edge_feats + e * dE);
  }
  return AC_OK;
}
