// Synthetic inputs for the matcher: Erdős–Rényi and R-MAT data graphs, N(0,1) features and planted patterns.
// The reference ships no runnable dataset (its rmat18/georgiy edge files are absent), so every workload is made
// here, deterministically from a seed.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <unordered_set>

#include "../../include/ac_mi355x.h"
#include "ac_host.hpp"

using namespace ac;

namespace {

// Sort packed (src << 32 | dst) keys in parallel (bucketed by the top bits), drop duplicates, and write them
// out sorted.  Returns the number of distinct keys.
int64_t sort_unique_keys(std::vector<uint64_t>& keys, int key_bits) {
  const int64_t n = (int64_t)keys.size();
  if (n == 0) return 0;
  const int nb_bits = 10;
  const int shift = key_bits > nb_bits ? key_bits - nb_bits : 0;
  const int64_t nb = (int64_t)1 << nb_bits;
  const int T = host_threads();
  std::vector<std::vector<int64_t>> counts(T, std::vector<int64_t>(nb, 0));
  parallel_chunks(n, [&](int64_t b, int64_t e, int t) {
    for (int64_t i = b; i < e; ++i) counts[t][std::min<uint64_t>(keys[i] >> shift, nb - 1)]++;
  }, T);
  std::vector<int64_t> bucket_off(nb + 1, 0);
  std::vector<std::vector<int64_t>> pos(T, std::vector<int64_t>(nb, 0));
  {
    int64_t acc = 0;
    for (int64_t k = 0; k < nb; ++k) {
      bucket_off[k] = acc;
      for (int t = 0; t < T; ++t) {
        pos[t][k] = acc;
        acc += counts[t][k];
      }
    }
    bucket_off[nb] = acc;
  }
  std::vector<uint64_t> tmp(n);
  parallel_chunks(n, [&](int64_t b, int64_t e, int t) {
    for (int64_t i = b; i < e; ++i) tmp[pos[t][std::min<uint64_t>(keys[i] >> shift, nb - 1)]++] = keys[i];
  }, T);
  std::vector<int64_t> uniq(nb, 0);
  parallel_items(nb, [&](int64_t k) {
    uint64_t* p = tmp.data() + bucket_off[k];
    int64_t len = bucket_off[k + 1] - bucket_off[k];
    std::sort(p, p + len);
    uniq[k] = std::unique(p, p + len) - p;
  }, T);
  int64_t out = 0;
  for (int64_t k = 0; k < nb; ++k) {
    std::memmove(keys.data() + out, tmp.data() + bucket_off[k], uniq[k] * sizeof(uint64_t));
    out += uniq[k];
  }
  keys.resize(out);
  return out;
}

}  // namespace

extern "C" int ac_gen_er(int64_t n, int64_t m, uint64_t seed, int32_t* srcs, int32_t* dsts) {
  if (n < 2 || m < 0 || m > n * (n - 1) || !srcs || !dsts || n > INT32_MAX) {
    set_global_error("ac_gen_er: need n >= 2, 0 <= m <= n(n-1), non-null outputs");
    return AC_ERR_ARG;
  }
  std::unordered_set<uint64_t> seen;
  seen.reserve((size_t)m * 2);
  std::vector<uint64_t> keys;
  keys.reserve(m);
  SeqRng rng(seed, kStreamEr);
  while ((int64_t)keys.size() < m) {
    uint64_t s = rng.below(n), d = rng.below(n);
    if (s == d) continue;
    uint64_t k = (s << 32) | d;
    if (seen.insert(k).second) keys.push_back(k);
  }
  std::sort(keys.begin(), keys.end());
  for (int64_t i = 0; i < m; ++i) {
    srcs[i] = (int32_t)(keys[i] >> 32);
    dsts[i] = (int32_t)(keys[i] & 0xffffffffu);
  }
  return AC_OK;
}

extern "C" int ac_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                           int64_t* num_edges, int32_t** srcs, int32_t** dsts) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0 || !num_edges ||
      !srcs || !dsts) {
    set_global_error("ac_gen_rmat: bad arguments");
    return AC_ERR_ARG;
  }
  const int64_t draws = (int64_t)edge_factor << scale;
  std::vector<uint64_t> keys(draws);
  const double ab = a + b, abc = a + b + c;
  // Each draw descends `scale` levels; level l picks a quadrant with one uniform number.
  parallel_chunks(draws, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; ++i) {
      uint64_t s = 0, d = 0;
      for (int l = 0; l < scale; ++l) {
        double u = unit_open(hash3(seed, kStreamRmat, (uint64_t)i * (uint64_t)scale + (uint64_t)l));
        uint64_t sb = u >= ab ? 1 : 0;
        uint64_t db = (u >= a && u < ab) || u >= abc ? 1 : 0;
        s = (s << 1) | sb;
        d = (d << 1) | db;
      }
      // Self-loops are marked with an all-ones key and sorted to the end.
      keys[i] = s == d ? ~0ull : ((s << scale) | d);
    }
  });
  int64_t u = sort_unique_keys(keys, 2 * scale);
  if (u > 0 && keys[u - 1] == ~0ull) --u;
  int32_t* S = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(u, 1));
  int32_t* D = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(u, 1));
  if (!S || !D) {
    std::free(S);
    std::free(D);
    set_global_error("ac_gen_rmat: out of host memory");
    return AC_ERR_OOM;
  }
  const uint64_t mask = ((uint64_t)1 << scale) - 1;
  parallel_chunks(u, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; ++i) {
      S[i] = (int32_t)(keys[i] >> scale);
      D[i] = (int32_t)(keys[i] & mask);
    }
  });
  *num_edges = u;
  *srcs = S;
  *dsts = D;
  return AC_OK;
}

extern "C" int ac_gen_normal(float* out, int64_t count, uint64_t seed, uint64_t stream, float sigma) {
  if (count < 0 || (count > 0 && !out)) {
    set_global_error("ac_gen_normal: bad arguments");
    return AC_ERR_ARG;
  }
  // Box–Muller on pairs (2j, 2j+1) of counter-based uniforms.
  const double two_pi = 6.283185307179586476925286766559;
  parallel_chunks((count + 1) / 2, [&](int64_t lo, int64_t hi, int) {
    for (int64_t j = lo; j < hi; ++j) {
      double u1 = unit_open(hash3(seed, stream, 2 * (uint64_t)j));
      double u2 = unit_open(hash3(seed, stream, 2 * (uint64_t)j + 1));
      double r = std::sqrt(-2.0 * std::log(u1));
      out[2 * j] = (float)(sigma * r * std::cos(two_pi * u2));
      if (2 * j + 1 < count) out[2 * j + 1] = (float)(sigma * r * std::sin(two_pi * u2));
    }
  });
  return AC_OK;
}

namespace {

// Incidence lists of the undirected view of a graph: for each vertex, the ids of the edges touching it.
struct Incidence {
  std::vector<int64_t> off, inc;
};

int build_incidence(int64_t n, int64_t m, const int32_t* srcs, const int32_t* dsts, Incidence& I, const char* who) {
  I.off.assign(n + 1, 0);
  for (int64_t e = 0; e < m; ++e) {
    if (srcs[e] < 0 || srcs[e] >= n || dsts[e] < 0 || dsts[e] >= n) {
      set_global_error(std::string(who) + ": edge id out of range");
      return AC_ERR_ARG;
    }
    I.off[srcs[e] + 1]++;
    if (dsts[e] != srcs[e]) I.off[dsts[e] + 1]++;
  }
  for (int64_t v = 0; v < n; ++v) I.off[v + 1] += I.off[v];
  I.inc.resize(I.off[n]);
  std::vector<int64_t> fill(I.off.begin(), I.off.end() - 1);
  for (int64_t e = 0; e < m; ++e) {
    I.inc[fill[srcs[e]]++] = e;
    if (dsts[e] != srcs[e]) I.inc[fill[dsts[e]]++] = e;
  }
  return AC_OK;
}

// Plants one pattern: a dense connected vertex set grown greedily from a random start, a BFS spanning tree of its
// induced edges plus random extra induced edges, then a random pattern-vertex numbering and edge order.
bool plant_one(int64_t n, const int32_t* srcs, const int32_t* dsts, const Incidence& I, int n_pv, int n_pe,
               uint64_t seed, int32_t* p_srcs, int32_t* p_dsts, int32_t* vmap, int64_t* emap) {
  const std::vector<int64_t>& off = I.off;
  const std::vector<int64_t>& inc = I.inc;
  SeqRng rng(seed, kStreamPattern);
  const int max_scan = 8192;  // incident edges examined per member (hubs are sampled from a random offset)
  for (int attempt = 0; attempt < 2000; ++attempt) {
    int64_t start = (int64_t)rng.below(n);
    if (off[start + 1] == off[start]) continue;
    std::vector<int64_t> members{start};
    std::unordered_map<int64_t, int> index{{start, 0}};
    bool stuck = false;
    while ((int)members.size() < n_pv) {
      // Score every outside neighbour by its number of incident edges into the current set and take the best
      // (ties broken by a hash), which grows dense, connected vertex sets.
      std::unordered_map<int64_t, int> score;
      for (int64_t u : members) {
        int64_t deg = off[u + 1] - off[u];
        int64_t take = std::min<int64_t>(deg, max_scan);
        int64_t base = deg > max_scan ? (int64_t)rng.below(deg) : 0;
        for (int64_t t = 0; t < take; ++t) {
          int64_t e = inc[off[u] + (base + t) % deg];
          int64_t w = srcs[e] == u ? dsts[e] : srcs[e];
          if (!index.count(w)) score[w]++;
        }
      }
      if (score.empty()) {
        stuck = true;
        break;
      }
      int64_t best = -1;
      int best_s = -1;
      uint64_t best_h = 0;
      for (auto& kv : score) {
        uint64_t h = hash3(seed, attempt, (uint64_t)kv.first);
        if (kv.second > best_s || (kv.second == best_s && h < best_h)) {
          best = kv.first;
          best_s = kv.second;
          best_h = h;
        }
      }
      index[best] = (int)members.size();
      members.push_back(best);
    }
    if (stuck) continue;
    // Induced edges (each data edge once), self-loops excluded.
    std::vector<int64_t> induced;
    {
      std::unordered_set<int64_t> seen;
      for (int64_t u : members)
        for (int64_t t = off[u]; t < off[u + 1]; ++t) {
          int64_t e = inc[t];
          if (srcs[e] != dsts[e] && index.count(srcs[e]) && index.count(dsts[e]) && seen.insert(e).second)
            induced.push_back(e);
        }
    }
    if ((int64_t)induced.size() < n_pe) continue;
    std::sort(induced.begin(), induced.end());
    // A spanning tree first (BFS over the induced undirected edges), then random extra induced edges.
    std::vector<char> in_tree(n_pv, 0), used(induced.size(), 0);
    std::vector<int64_t> chosen;
    std::vector<int> queue{0};
    in_tree[0] = 1;
    for (size_t h = 0; h < queue.size(); ++h) {
      int u = queue[h];
      for (size_t i = 0; i < induced.size(); ++i) {
        if (used[i]) continue;
        int a = index[srcs[induced[i]]], b = index[dsts[induced[i]]];
        int w = a == u ? b : (b == u ? a : -1);
        if (w < 0 || in_tree[w]) continue;
        in_tree[w] = 1;
        used[i] = 1;
        chosen.push_back(induced[i]);
        queue.push_back(w);
      }
    }
    if ((int)chosen.size() != n_pv - 1) continue;
    std::vector<int64_t> rest;
    for (size_t i = 0; i < induced.size(); ++i)
      if (!used[i]) rest.push_back(induced[i]);
    for (int64_t i = (int64_t)rest.size() - 1; i > 0; --i) std::swap(rest[i], rest[rng.below(i + 1)]);
    for (int64_t i = 0; (int64_t)chosen.size() < n_pe; ++i) chosen.push_back(rest[i]);
    // Random pattern-vertex numbering and pattern-edge order.
    std::vector<int> perm(n_pv);
    for (int i = 0; i < n_pv; ++i) perm[i] = i;
    for (int i = n_pv - 1; i > 0; --i) std::swap(perm[i], perm[rng.below(i + 1)]);
    for (int64_t i = (int64_t)chosen.size() - 1; i > 0; --i) std::swap(chosen[i], chosen[rng.below(i + 1)]);
    for (int i = 0; i < n_pv; ++i) vmap[perm[i]] = (int32_t)members[i];
    for (int q = 0; q < n_pe; ++q) {
      int64_t e = chosen[q];
      p_srcs[q] = perm[index[srcs[e]]];
      p_dsts[q] = perm[index[dsts[e]]];
      emap[q] = e;
    }
    return true;
  }
  return false;
}

}  // namespace

extern "C" int ac_gen_pattern(int64_t n, int64_t m, const int32_t* srcs, const int32_t* dsts, int n_pv, int n_pe,
                              uint64_t seed, int32_t* p_srcs, int32_t* p_dsts, int32_t* vmap, int64_t* emap) {
  if (n < 1 || m < 1 || !srcs || !dsts || n_pv < 1 || n_pe < n_pv - 1 || !p_srcs || !p_dsts || !vmap || !emap ||
      n_pv > n) {
    set_global_error("ac_gen_pattern: bad arguments (need n_pe >= n_pv - 1)");
    return AC_ERR_ARG;
  }
  Incidence I;
  int rc = build_incidence(n, m, srcs, dsts, I, "ac_gen_pattern");
  if (rc != AC_OK) return rc;
  if (plant_one(n, srcs, dsts, I, n_pv, n_pe, seed, p_srcs, p_dsts, vmap, emap)) return AC_OK;
  set_global_error("ac_gen_pattern: no connected subgraph with that many induced edges was found");
  return AC_ERR_ARG;
}

extern "C" int ac_gen_patterns(int64_t n, int64_t m, const int32_t* srcs, const int32_t* dsts, int k,
                               const int32_t* n_pv, const int32_t* n_pe, const uint64_t* seeds, int32_t* p_srcs,
                               int32_t* p_dsts, int32_t* vmap, int64_t* emap, int32_t* status) {
  if (n < 1 || m < 1 || !srcs || !dsts || k < 1 || !n_pv || !n_pe || !seeds || !p_srcs || !p_dsts || !vmap ||
      !emap || !status) {
    set_global_error("ac_gen_patterns: bad arguments");
    return AC_ERR_ARG;
  }
  std::vector<int64_t> voff(k + 1, 0), eoff(k + 1, 0);
  for (int i = 0; i < k; ++i) {
    if (n_pv[i] < 1 || n_pv[i] > n || n_pe[i] < n_pv[i] - 1) {
      set_global_error("ac_gen_patterns: pattern " + std::to_string(i) + " needs 1 <= vertices <= n and edges >= vertices - 1");
      return AC_ERR_ARG;
    }
    voff[i + 1] = voff[i] + n_pv[i];
    eoff[i + 1] = eoff[i] + n_pe[i];
  }
  Incidence I;
  int rc = build_incidence(n, m, srcs, dsts, I, "ac_gen_patterns");
  if (rc != AC_OK) return rc;
  parallel_items(k, [&](int64_t i) {
    status[i] = plant_one(n, srcs, dsts, I, n_pv[i], n_pe[i], seeds[i], p_srcs + eoff[i], p_dsts + eoff[i],
                          vmap + voff[i], emap + eoff[i])
                    ? 0
                    : 1;
  });
  return AC_OK;
}

extern "C" void ac_free(void* p) { std::free(p); }
