// Host-side code under AddressSanitizer (test infrastructure): the library's host-only translation units (file formats
// and generators, application_classification_amd/csrc/ac_io.cpp and ac_gen.cpp) and the oracle
// (oracle/ac_oracle.cpp), built with g++ -fsanitize=address by tests/test_host_asan.py, which runs this driver.  It
// walks the paths that allocate, parse or index: CSV and .bin readers and writers at both precisions (good and
// malformed files), the generators, pattern planting, and the oracle on a planted pattern.  Exit status 0 and no
// sanitizer report means clean.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/ac_mi355x.h"

extern "C" {
typedef struct or_graph {
  int64_t nV, dV, nE, dE;
  const double* vfeat;
  const double* efeat;
  const int64_t* src;
  const int64_t* dst;
} or_graph;
int or_run_f64(const or_graph* D, const or_graph* P, int n_iter, unsigned flags, int threads, double* mu_out,
               double* fmax_out, double* rmax_out, double* elapsed_s);
}

static int fails = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      ++fails;                                                    \
    }                                                             \
  } while (0)

static void write_file(const std::string& path, const std::string& text) {
  FILE* f = std::fopen(path.c_str(), "wb");
  std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string ref = argv[1], tmp = argv[2];

  // Reference CSVs at both precisions, .bin round trips.
  ac_graph64 g64;
  CHECK(ac_graph64_load_csv((ref + "/georgiyPattern.Vertex.csv").c_str(), (ref + "/georgiyPattern.Edges.csv").c_str(),
                            &g64) == AC_OK);
  CHECK(g64.num_nodes == 10 && g64.edge_feat_dim == 17);
  CHECK(ac_graph64_save_bin((tmp + "/p.bin").c_str(), &g64) == AC_OK);
  ac_graph64 h64;
  CHECK(ac_graph64_load_bin((tmp + "/p.bin").c_str(), &h64) == AC_OK);
  CHECK(std::memcmp(g64.edge_feats, h64.edge_feats, sizeof(double) * g64.num_edges * g64.edge_feat_dim) == 0);
  ac_graph64_free(&h64);
  ac_host_graph g32;
  CHECK(ac_graph_load_bin((tmp + "/p.bin").c_str(), &g32) == AC_OK);
  CHECK(ac_graph_save_csv((tmp + "/v.csv").c_str(), (tmp + "/e.csv").c_str(), &g32) == AC_OK);
  ac_host_graph c32;
  CHECK(ac_graph_load_csv((tmp + "/v.csv").c_str(), (tmp + "/e.csv").c_str(), &c32) == AC_OK);
  CHECK(c32.num_edges == g32.num_edges);
  ac_graph_free(&c32);
  ac_graph_free(&g32);
  ac_graph64_free(&g64);

  // Malformed inputs: truncated .bin, ragged and out-of-range CSV rows, odd tokens, empty files.
  write_file(tmp + "/t.bin", std::string("\x02\x00\x00\x00\x00\x00\x00\x00", 8));
  CHECK(ac_graph64_load_bin((tmp + "/t.bin").c_str(), &h64) == AC_ERR_IO);
  CHECK(ac_graph_load_bin((tmp + "/t.bin").c_str(), &g32) == AC_ERR_IO);
  write_file(tmp + "/bv.csv", "2 3\r\n0 1.5 -2e3\r\n1 .5 7.\r\n");
  write_file(tmp + "/be.csv", "2 4\r\n0 1 1e-320 x\r\n1 9 0.5 0.25\r\n");
  CHECK(ac_graph64_load_csv((tmp + "/bv.csv").c_str(), (tmp + "/be.csv").c_str(), &h64) == AC_ERR_IO);
  write_file(tmp + "/be.csv", "2 4\r\n0 1 0.1\r\n1 0 0.2 0.3\r\n");  // ragged
  CHECK(ac_graph64_load_csv((tmp + "/bv.csv").c_str(), (tmp + "/be.csv").c_str(), &h64) == AC_ERR_IO);
  write_file(tmp + "/be.csv", "");
  CHECK(ac_graph64_load_csv((tmp + "/bv.csv").c_str(), (tmp + "/be.csv").c_str(), &h64) == AC_OK);
  CHECK(h64.num_nodes == 2 && h64.num_edges == 0 && h64.node_feats[1] == -2000.0);
  ac_graph64_free(&h64);
  CHECK(ac_graph_load_csv((tmp + "/none.csv").c_str(), (tmp + "/none.csv").c_str(), &g32) == AC_ERR_IO);

  // Generators and planting.
  const int64_t n = 1 << 10;
  int64_t m = 0;
  int32_t *s = nullptr, *d = nullptr;
  CHECK(ac_gen_rmat(10, 8, 0.57, 0.19, 0.19, 3, &m, &s, &d) == AC_OK && m > 0);
  std::vector<int32_t> es(4000), ed(4000);
  CHECK(ac_gen_er(n, 4000, 5, es.data(), ed.data()) == AC_OK);
  std::vector<float> vf((size_t)n * 8), ef((size_t)m * 8);
  CHECK(ac_gen_normal(vf.data(), (int64_t)vf.size(), 1, 10, 1.f) == AC_OK);
  CHECK(ac_gen_normal(ef.data(), (int64_t)ef.size(), 1, 11, 1.f) == AC_OK);
  const int pv = 6, pe = 8;
  std::vector<int32_t> ps(pe), pd(pe), vmap(pv);
  std::vector<int64_t> emap(pe);
  CHECK(ac_gen_pattern(n, m, s, d, pv, pe, 7, ps.data(), pd.data(), vmap.data(), emap.data()) == AC_OK);
  const int k = 3;
  const int32_t npv[k] = {4, 5, 40}, npe[k] = {4, 6, 400};
  const uint64_t seeds[k] = {1, 2, 3};
  std::vector<int32_t> kps(410), kpd(410), kvm(49), st(k);
  std::vector<int64_t> kem(410);
  CHECK(ac_gen_patterns(n, m, s, d, k, npv, npe, seeds, kps.data(), kpd.data(), kvm.data(), kem.data(), st.data()) ==
        AC_OK);

  // The oracle on the planted pattern (features copied, so the plant is recovered).
  std::vector<double> dv(vf.begin(), vf.end()), de(ef.begin(), ef.end()), pvf((size_t)pv * 8), pef((size_t)pe * 8);
  std::vector<int64_t> ds(s, s + m), dd(d, d + m), qs(ps.begin(), ps.end()), qd(pd.begin(), pd.end());
  for (int p = 0; p < pv; ++p)
    for (int c = 0; c < 8; ++c) pvf[(size_t)p * 8 + c] = dv[(size_t)vmap[p] * 8 + c];
  for (int q = 0; q < pe; ++q)
    for (int c = 0; c < 8; ++c) pef[(size_t)q * 8 + c] = de[(size_t)emap[q] * 8 + c];
  or_graph D{n, 8, m, 8, dv.data(), de.data(), ds.data(), dd.data()};
  or_graph P{pv, 8, pe, 8, pvf.data(), pef.data(), qs.data(), qd.data()};
  std::vector<double> mu((size_t)n * pv);
  for (unsigned flags : {0u, 1u, 2u, 4u}) CHECK(or_run_f64(&D, &P, -1, flags, 2, mu.data(), nullptr, nullptr, nullptr) == 0);
  for (int p = 0; p < pv; ++p) {
    int64_t best = 0;
    for (int64_t v = 1; v < n; ++v)
      if (mu[(size_t)v * pv + p] > mu[(size_t)best * pv + p]) best = v;
    CHECK(best == vmap[p]);
  }
  ac_free(s);
  ac_free(d);
  std::printf("asan driver: %d failed checks\n", fails);
  return fails ? 1 : 0;
}
