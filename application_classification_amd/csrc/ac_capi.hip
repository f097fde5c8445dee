// C-ABI implementation: the matcher context, graph upload and the driver's prep + loop (src/main.cu:66-257).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/ac_mi355x.h"
#include "ac_adjoint.hpp"
#include "ac_host.hpp"
#include "ac_internal.hpp"

using ac::Lse;
using ac::Lse2;
using ac::QPart;
using ac::QStat;
using ac::IsoCol;
using ac::Bundle;
using ac::Run;
using ac::Split;

namespace {

// rocTX range over a scope, named like the reference's NVTX ranges (src/main.cu:107,161-244: "prep", "loop" per
// iteration, "step1".."step5"); the fused kernels each cover several of the reference's steps, and their ranges name
// all of them.  The ranges time the host-side enqueue (launches are asynchronous) and appear under rocprofv3
// --marker-trace.
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

// Edges per run of an edge walk (a tuning knob, AC_RUN_EDGES).  Vertices whose edge list is longer are split into
// pieces of this length.
int run_edges() {
  if (const char* e = std::getenv("AC_RUN_EDGES")) {
    const int v = std::atoi(e);
    if (v >= 4 && v <= 4096) return v;
  }
  return 128;  // measured at R-MAT 22 against 48-512 (with walk batch 4)
}

int round8(int x) { return (x + 7) / 8 * 8; }

// Blocks of a vertex pass.  Each block walks its tiles grid-stride, and passes A and C leave one partial per pattern
// edge / vertex that is merged serially afterwards: one round of resident blocks (no partly filled last round, few
// partials) is fastest for them at R-MAT scale 22; the output passes, which leave no partials, use two rounds.
int vertex_blocks(const char* env, int resident, int rounds, int64_t ntiles) {
  if (const char* e = std::getenv(env)) {  // tuning override
    const int v = std::atoi(e);
    if (v >= 64 && v <= 65536) return (int)std::min<int64_t>(v, std::max<int64_t>(ntiles, 1));
  }
  const int64_t want = resident > 0 ? (int64_t)resident * rounds : 1024;
  return (int)std::max<int64_t>(1, std::min(want, ntiles));
}

// Row stride of the gathered VR rows: whole 128-byte lines (or an exact fraction of one), so that a gather of one
// row never straddles two lines.
int gather_stride(int Qs) {
  if (Qs > 32) return (Qs + 31) / 32 * 32;
  int s = 4;
  while (s < Qs) s *= 2;
  return s;
}

// Largest single device allocation a context makes (AC_MAX_ALLOC_BYTES, read at each allocation; unset: no cap).
// A larger request fails as out-of-memory, the way an allocation too large for HBM does: for sharing a GPU with other
// work, and for exercising the out-of-memory path.
size_t max_alloc_bytes() {
  const char* e = std::getenv("AC_MAX_ALLOC_BYTES");
  if (!e || !*e) return SIZE_MAX;
  const unsigned long long v = std::strtoull(e, nullptr, 10);
  return v ? (size_t)v : SIZE_MAX;
}

// Device buffer that only grows: a smaller request reuses the allocation (n is the logical size), so runs of
// patterns of different sizes do not free and re-allocate multi-GB buffers.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = cap = 0;
  }
  hipError_t alloc(size_t count) {
    if (count <= cap && p) {
      n = count;
      return hipSuccess;
    }
    release();
    if (count == 0) return hipSuccess;
    if (count > max_alloc_bytes() / sizeof(T)) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = cap = count;
    return hipSuccess;
  }
  hipError_t upload(const T* host, size_t count, hipStream_t st) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    e = hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
  }
};

// Runs over n vertices whose edge lists are the ranges [off[t], off[t + 1]) in sequence: vertex vid[t] (t itself
// when vid is null).  consecutive: a run may only join vertices whose ids follow each other (the CSC walk steps its
// destination by one at each new vertex); the CSR walk reads each edge's source and takes any order.
void build_runs(const std::vector<int64_t>& off, int64_t n, std::vector<Run>& runs, std::vector<Split>& splits,
                int& nslots, const int* vid = nullptr, bool consecutive = true) {
  const int kRunEdges = run_edges();
  runs.clear();
  splits.clear();
  nslots = 0;
  bool open = false;
  int64_t b = 0, len = 0, last_v = -2;
  auto close = [&](int64_t e) {
    if (open) runs.push_back(Run{(int)b, (int)e, -1, 0});
    open = false;
  };
  for (int64_t t = 0; t < n; ++t) {
    const int64_t v = vid ? vid[t] : t;
    const int64_t deg = off[t + 1] - off[t];
    if (deg == 0) continue;
    if (deg > kRunEdges) {
      close(off[t]);
      const int pieces = (int)((deg + kRunEdges - 1) / kRunEdges);
      splits.push_back(Split{(int)v, nslots, pieces, 0});
      for (int i = 0; i < pieces; ++i) {
        const int64_t kb = off[t] + (int64_t)i * kRunEdges;
        runs.push_back(Run{(int)kb, (int)std::min(off[t + 1], kb + kRunEdges), nslots++, 0});
      }
      continue;
    }
    if (open && (len + deg > kRunEdges || (consecutive && v != last_v + 1))) close(off[t]);
    last_v = v;
    if (!open) {
      open = true;
      b = off[t];
      len = 0;
    }
    len += deg;
  }
  close(off[n]);
}

}  // namespace

struct ac_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // A second stream (default priority, like the first) for prep work that does not depend on the edge costs or
  // that only reads them beside another kernel; xev fork and join it (no timing).
  hipStream_t aux = nullptr;
  hipEvent_t xev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::string err;

  // Data graph (src/main.cu:8-19 / 49-62) plus its CSC and CSR views.
  bool has_data = false;
  int64_t dn = 0, dV = 0, de = 0, dE = 0;
  int64_t max_outdeg = 0;  // the largest out-degree (sets the headroom of the adjoint's fixed-point sums)
  DevBuf<float> vfeat, efeat;
  DevBuf<int2> csc_sd, csr_ks;
  DevBuf<int> csc_eid;
  DevBuf<Run> runs_in, runs_out;
  DevBuf<Split> splits_in, splits_out;
  // The plain walk's fixup work: chunks of at most kFixChunk pieces of a split vertex ({v, first piece, count, out
  // partial or -1}) and the vertices of several chunks ({v, first partial, count}).
  DevBuf<Split> fix_chunks, fix_multi;
  int nfix_chunks = 0, nfix_multi = 0, nfix_parts = 0;
  std::vector<Run> h_runs_in;
  std::vector<int> h_run_src0;  // first (least) source of each in-edge run
  // Internal vertex order: vertices with in- and out-edges, in-edges only, out-edges only, then isolated ones, each
  // class in ascending caller order.  perm maps an internal vertex to the caller's; the vertex passes stream the
  // first n_act = dn - n_iso rows only.
  DevBuf<int> d_perm;
  int64_t n_act = 0, n_iso = 0, n_both = 0, n_inonly = 0;
  int64_t nruns_in = 0, nruns_out = 0;
  int nsplit_in = 0, nsplit_out = 0, nslots_in = 0, nslots_out = 0;

  // Per-dimension cdist weights (the adjoint's learnable feature metric); absent = the reference's cdist.
  bool has_wv = false, has_we = false;
  DevBuf<double> wv, we;
  std::vector<double> h_wv, h_we;

  // Pattern graph.
  bool has_patt = false;
  int P = 0, Q = 0, Qs = 0, pdV = 0, pdE = 0;
  std::vector<float> pvfeat, pefeat;
  std::vector<int> ps, pd;
  DevBuf<double> pf_v, pf_e;
  DevBuf<int> d_ps, d_pd, d_inc_off, d_inc;

  // Work buffers.
  DevBuf<float> CE, CV, FMaxN, VFm, minCEo, lsO, VR, OUT;
  // Bundle layouts of the in-edge runs, one per number of runs per bundle G: G follows the pattern through Qs,
  // so a few are kept (most recently used last) for runs of patterns of different widths.  Each holds the rows' edge
  // endpoints and edge ids and the CSR walk's {row, src} pairs.
  struct BundleSet {
    int G = 0;
    DevBuf<Bundle> bundles;
    DevBuf<int> gslot, bsrc, bdst, beid;
    DevBuf<int2> csr_kb;
    int64_t nbundles = 0, brows = 0;
  };
  std::vector<std::unique_ptr<BundleSet>> bsets;
  BundleSet* bs = nullptr;
  int walk_B = 0;          // steps per batch of the edge walk (0: default)
  bool use_graph = false;  // replay iteration pairs from a captured graph
  bool prune = true;       // skip bundles whose edges cannot raise FMax above the fill (exact)
  bool pack_costs = true;  // walks read 24-bit codes of the costs where a bundle allows them (exact)
  // The PNNL Cnull term (off: the reference, src/main.cu:83).  cnull_groups: the pattern-edge ranges over which it is
  // normalised, one per pattern (a batched slab holds several); d_cnull holds it per pattern edge when on.
  bool cnull = false;
  std::vector<int> cnull_groups;
  DevBuf<double> d_cnull;
  bool costs_packed = false;  // the last prep wrote CEp / bpack for the current CE
  DevBuf<unsigned> CEp;    // [rows * Qs * 3 / 4] packed costs of the bundles whose base is nonzero
  DevBuf<unsigned> bpack;  // [bundles] base of a bundle's codes, 0 = FP32 rows
  bool use_mfma_cdist = true;  // feature distances on FP64 matrix cores (GEMM form, cancellation-guarded)
  DevBuf<double> cd_scratch;   // the matrix-core operands of the pattern features
  DevBuf<float> bmin;      // [bundles][Qs] least cost per bundle and pattern edge
  DevBuf<double> carry_in, carry_out;
  DevBuf<float> fix_part;  // [nfix_parts][Qs]
  DevBuf<QPart> qpart;
  DevBuf<Lse> ppart, lpart;
  DevBuf<QStat> qstat;  // [3][Q]: prep state, then by iteration parity
  DevBuf<Lse2> pstat, lse_cv, lse_ce;
  DevBuf<IsoCol> iso;   // [P] the isolated vertices' columns of CV (when n_iso > 0)
  int nb_vertex = 0, nb_vertex_c = 0, nb_vertex_o = 0, nb_col = 0, nb_src = 0, VT = 0, VRs = 0;
  int64_t work_key[5] = {-1, -1, -1, -1, -1};

  bool have_result = false;
  int last_iters = 0;

  // Adjoint tape: with tape_K > 0 a run records the argmax rows / vertices its max stages used and checkpoints the
  // iteration state every tape_K iterations; ac_backward recomputes each segment from its checkpoint.  With
  // tape_K = 1 every state is a checkpoint (the run writes each into its slot, the last one included) and the run's
  // walks also record their argmax rows (tapeFA, one [dn][Qs] slot per iteration), so ac_backward recomputes nothing.
  int tape_K = 0;
  bool tape_valid = false;
  int tape_iters = 0;
  DevBuf<int> FA0, AO, am0, carg_in, tapeFA;
  DevBuf<float> ckF, ckV, sgF, sgV;
  DevBuf<QStat> ckQ, sgQ;
  DevBuf<Lse2> ckP, sgP;
  // Adjoint work buffers.
  DevBuf<float> gMUin, gMU, gCV, gEF, gER, gVR, gVF, gCE, apv;
  DevBuf<unsigned long long> gVRfix;  // per (vertex, pattern edge), fixed-point sums of a step's FMax edge terms
  DevBuf<unsigned> gmax;              // per step back, the largest |gE_F| (float bits)
  DevBuf<double> vpart;               // partial column sums of gMU from the vertex kernel
  DevBuf<float> gERsum;               // per (vertex, pattern edge), gE_R summed over the steps back
  DevBuf<double> cD, colsum, bpart, red4, cq, opart, osum, MUd;
  DevBuf<int> api;

  // Batched scoring: the patterns (host copies, concatenated), where each one's result lives in `bout`, and the
  // statistics of the last batched run.
  struct BatchPattern {
    int P, Q;
    int64_t v0, e0, out_off;
  };
  std::vector<BatchPattern> bp;
  std::vector<float> b_vfeat, b_efeat;
  std::vector<int> b_src, b_dst;
  int64_t b_dV = 0, b_dE = 0;
  int slab_edges = 128;
  DevBuf<float> bout;
  bool have_batch = false;
  int64_t b_info[4] = {0, 0, 0, 0};
  double b_prep_ms = 0.0, b_loop_ms = 0.0;

  // Profiling: one event after every launch, tagged with the stage it closes.
  bool profiling = false;
  std::vector<hipEvent_t> pev;
  std::vector<int> pev_stage;
  size_t pev_used = 0;

  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
};

#define HIPCHK(ctx, expr)                                                                              \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess) return (ctx)->fail(_e == hipErrorOutOfMemory ? AC_ERR_OOM : AC_ERR_HIP,      \
                                             std::string(#expr) + ": " + hipGetErrorString(_e));       \
  } while (0)

extern "C" int ac_create(ac_ctx** out, int device) {
  if (!out) {
    ac::set_global_error("ac_create: null output pointer");
    return AC_ERR_ARG;
  }
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    ac::set_global_error("ac_create: no HIP device is visible");
    return AC_ERR_NODEV;
  }
  if (device < 0 || device >= count) {
    ac::set_global_error("ac_create: device " + std::to_string(device) + " does not exist");
    return AC_ERR_NODEV;
  }
  if (hipSetDevice(device) != hipSuccess) {
    ac::set_global_error("ac_create: hipSetDevice failed");
    return AC_ERR_HIP;
  }
  ac_ctx* c = new ac_ctx();
  c->device = device;
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) == hipSuccess;
  for (auto& e : c->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  for (auto& e : c->xev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    for (auto& e : c->ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : c->xev)
      if (e) (void)hipEventDestroy(e);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    ac::set_global_error("ac_create: could not create the stream / events");
    return AC_ERR_HIP;
  }
  *out = c;
  return AC_OK;
}

extern "C" int ac_destroy(ac_ctx* ctx) {
  if (!ctx) return AC_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->aux) (void)hipStreamSynchronize(ctx->aux);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->xev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->pev) (void)hipEventDestroy(e);
  if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return AC_OK;
}

extern "C" const char* ac_last_error(const ac_ctx* ctx) { return ctx ? ctx->err.c_str() : ac::global_error(); }

extern "C" void* ac_stream(ac_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

static int set_data(ac_ctx* ctx, int64_t nV, int64_t dV, const float* vf, int64_t nE, int64_t dE, const float* ef,
                    const int32_t* src, const int32_t* dst) {
  if (nE >= (int64_t)INT32_MAX - 4096)
    return ctx->fail(AC_ERR_ARG, "ac_set_graph: more than 2^31 edges is not supported");
  // Internal order by degree class (see ac_ctx::d_perm); every structure below uses internal vertex ids.
  std::vector<int> perm, inv(nV);
  int64_t n_iso = 0, n_both = 0, n_inonly = 0;
  std::vector<int32_t> isrc(nE), idst(nE);
  {
    std::vector<uint8_t> cls(nV, 3);  // bit 0 clear: has in-edges; bit 1 clear: has out-edges
    for (int64_t e = 0; e < nE; ++e) {
      cls[dst[e]] &= 2;
      cls[src[e]] &= 1;
    }
    // Class key: 0 both, 1 in-edges only (cls 2), 2 out-edges only (cls 1), 3 isolated.
    const uint8_t key[4] = {0, 2, 1, 3};
    perm.reserve(nV);
    for (int c = 0; c < 4; ++c)
      for (int64_t v = 0; v < nV; ++v)
        if (key[cls[v]] == c) {
          inv[v] = (int)perm.size();
          perm.push_back((int)v);
          if (c == 3) ++n_iso;
          if (c == 0) ++n_both;
          if (c == 1) ++n_inonly;
        }
    for (int64_t e = 0; e < nE; ++e) {
      isrc[e] = inv[src[e]];
      idst[e] = inv[dst[e]];
    }
  }
  src = isrc.data();
  dst = idst.data();
  // CSR (by source, stable) and CSC (by destination, then source) views.  The CSC is filled from the edges in CSR
  // order, so each vertex's in-edges ascend by source: a piece of a long in-edge list then gathers from one band of
  // sources (the maxima do not depend on the order).
  std::vector<int64_t> in_off(nV + 1, 0), out_off(nV + 1, 0);
  for (int64_t e = 0; e < nE; ++e) {
    in_off[dst[e] + 1]++;
    out_off[src[e] + 1]++;
  }
  for (int64_t v = 0; v < nV; ++v) {
    in_off[v + 1] += in_off[v];
    out_off[v + 1] += out_off[v];
  }
  std::vector<int2> csc_sd(nE), csr_ks(nE);
  std::vector<int> csc_eid(nE), csc_pos(nE), by_src(nE);
  {
    std::vector<int64_t> fill(out_off.begin(), out_off.end() - 1);
    for (int64_t e = 0; e < nE; ++e) by_src[fill[src[e]]++] = (int)e;
  }
  {
    std::vector<int64_t> fill(in_off.begin(), in_off.end() - 1);
    for (int64_t i = 0; i < nE; ++i) {
      const int64_t e = by_src[i];
      const int64_t k = fill[dst[e]]++;
      csc_eid[k] = (int)e;
      csc_sd[k] = make_int2(src[e], dst[e]);
      csc_pos[e] = (int)k;
    }
  }
  // The CSR pairs list the sources by decreasing out-degree (ties by id), each source's edges in CSR order.  The
  // per-source walk (k_walk_src) gives a run to each group of lanes and finishes a source wave-wide whenever any
  // group's source changes; with the runs of a wave made of sources of one degree, the groups change source on the
  // same steps.
  std::vector<int> sord;
  std::vector<int64_t> sord_off(1, 0);
  int64_t max_outdeg = 0;
  {
    for (int64_t v = 0; v < nV; ++v) {
      if (out_off[v + 1] > out_off[v]) sord.push_back((int)v);
      max_outdeg = std::max(max_outdeg, out_off[v + 1] - out_off[v]);
    }
    std::stable_sort(sord.begin(), sord.end(), [&](int a, int b) {
      return out_off[a + 1] - out_off[a] > out_off[b + 1] - out_off[b];
    });
    sord_off.reserve(sord.size() + 1);
    int64_t i = 0;
    for (const int v : sord) {
      for (int64_t k = out_off[v]; k < out_off[v + 1]; ++k, ++i)
        csr_ks[i] = make_int2(csc_pos[by_src[k]], src[by_src[k]]);
      sord_off.push_back(i);
    }
  }
  std::vector<Run> rin, rout;
  std::vector<Split> sin, sout;
  int slin = 0, slout = 0;
  build_runs(in_off, nV, rin, sin, slin);
  build_runs(sord_off, (int64_t)sord.size(), rout, sout, slout, sord.data(), false);
  std::vector<Split> fch, fms;
  int nparts = 0;
  for (const Split& S : sin) {
    const int nch = (S.count + ac::kFixChunk - 1) / ac::kFixChunk;
    if (nch == 1) {
      fch.push_back(Split{S.v, S.first, S.count, -1});
      continue;
    }
    fms.push_back(Split{S.v, nparts, nch, 0});
    for (int c = 0; c < nch; ++c)
      fch.push_back(Split{S.v, S.first + c * ac::kFixChunk, std::min(ac::kFixChunk, S.count - c * ac::kFixChunk),
                          nparts + c});
    nparts += nch;
  }

  ctx->has_data = false;
  hipStream_t st = ctx->stream;
  HIPCHK(ctx, ctx->vfeat.upload(vf, (size_t)(nV * dV), st));
  HIPCHK(ctx, ctx->d_perm.upload(perm.data(), perm.size(), st));
  HIPCHK(ctx, ctx->efeat.upload(ef, (size_t)(nE * dE), st));
  HIPCHK(ctx, ctx->csc_sd.upload(csc_sd.data(), csc_sd.size(), st));
  HIPCHK(ctx, ctx->csr_ks.upload(csr_ks.data(), csr_ks.size(), st));
  HIPCHK(ctx, ctx->csc_eid.upload(csc_eid.data(), csc_eid.size(), st));
  HIPCHK(ctx, ctx->runs_in.upload(rin.data(), rin.size(), st));
  HIPCHK(ctx, ctx->runs_out.upload(rout.data(), rout.size(), st));
  HIPCHK(ctx, ctx->splits_in.upload(sin.data(), sin.size(), st));
  HIPCHK(ctx, ctx->splits_out.upload(sout.data(), sout.size(), st));
  HIPCHK(ctx, ctx->fix_chunks.upload(fch.data(), fch.size(), st));
  HIPCHK(ctx, ctx->fix_multi.upload(fms.data(), fms.size(), st));
  ctx->nfix_chunks = (int)fch.size();
  ctx->nfix_multi = (int)fms.size();
  ctx->nfix_parts = nparts;
  ctx->dn = nV;
  ctx->max_outdeg = max_outdeg;
  ctx->n_iso = n_iso;
  ctx->n_both = n_both;
  ctx->n_inonly = n_inonly;
  ctx->n_act = nV - n_iso;
  ctx->dV = dV;
  ctx->de = nE;
  ctx->dE = dE;
  ctx->nruns_in = (int64_t)rin.size();
  ctx->nruns_out = (int64_t)rout.size();
  ctx->nsplit_in = (int)sin.size();
  ctx->nsplit_out = (int)sout.size();
  ctx->nslots_in = slin;
  ctx->nslots_out = slout;
  ctx->h_run_src0.resize(rin.size());
  for (size_t i = 0; i < rin.size(); ++i) ctx->h_run_src0[i] = csc_sd[rin[i].kb].x;
  ctx->h_runs_in = std::move(rin);
  ctx->bsets.clear();  // bundle layouts follow the graph, not only its sizes
  ctx->bs = nullptr;
  ctx->work_key[0] = -1;
  ctx->has_data = true;
  ctx->have_result = false;
  return AC_OK;
}

static int set_pattern(ac_ctx* ctx, int64_t nV, int64_t dV, const float* vf, int64_t nE, int64_t dE, const float* ef,
                       const int32_t* src, const int32_t* dst) {
  if (nV > 256 || nE > 256)
    return ctx->fail(AC_ERR_ARG, "ac_set_graph: patterns are limited to 256 vertices and 256 edges");
  ctx->has_patt = false;
  ctx->P = (int)nV;
  ctx->Q = (int)nE;
  ctx->Qs = (int)((nE + 3) / 4 * 4);
  if (ctx->Qs == 0) ctx->Qs = 4;
  ctx->pdV = (int)dV;
  ctx->pdE = (int)dE;
  ctx->pvfeat.assign(vf, vf + nV * dV);
  ctx->pefeat.assign(ef, ef + nE * dE);
  ctx->ps.assign(src, src + nE);
  ctx->pd.assign(dst, dst + nE);
  // Pattern features in FP64 as [count rounded up to 8 / 8][dim][8] with zero padding: cdist reads the eight
  // patterns of a group together, one feature at a time.
  const int Pp = round8((int)nV), Qp = round8((int)nE);
  std::vector<double> pv((size_t)dV * Pp, 0.0), pe((size_t)dE * Qp, 0.0);
  for (int64_t p = 0; p < nV; ++p)
    for (int64_t k = 0; k < dV; ++k) pv[((p / 8) * dV + k) * 8 + p % 8] = vf[p * dV + k];
  for (int64_t q = 0; q < nE; ++q)
    for (int64_t k = 0; k < dE; ++k) pe[((q / 8) * dE + k) * 8 + q % 8] = ef[q * dE + k];
  // Incidence of each pattern vertex in ascending q: the FMax term (q ends at p) before the RMax term (q starts at p).
  std::vector<int> inc_off(nV + 1, 0), inc;
  for (int64_t p = 0; p < nV; ++p) {
    inc_off[p] = (int)inc.size();
    for (int64_t q = 0; q < nE; ++q) {
      if (dst[q] == p) inc.push_back((int)(q << 1));
      if (src[q] == p) inc.push_back((int)(q << 1) | 1);
    }
  }
  inc_off[nV] = (int)inc.size();
  if (inc.empty()) inc.push_back(0);
  hipStream_t st = ctx->stream;
  HIPCHK(ctx, ctx->pf_v.upload(pv.data(), pv.size(), st));
  HIPCHK(ctx, ctx->pf_e.upload(pe.data(), pe.size(), st));
  std::vector<int> ps_pad(ctx->ps), pd_pad(ctx->pd);
  if (ps_pad.empty()) ps_pad.push_back(0), pd_pad.push_back(0);
  HIPCHK(ctx, ctx->d_ps.upload(ps_pad.data(), ps_pad.size(), st));
  HIPCHK(ctx, ctx->d_pd.upload(pd_pad.data(), pd_pad.size(), st));
  HIPCHK(ctx, ctx->d_inc_off.upload(inc_off.data(), inc_off.size(), st));
  HIPCHK(ctx, ctx->d_inc.upload(inc.data(), inc.size(), st));
  ctx->cnull_groups.assign({0, (int)nE});
  ctx->has_patt = true;
  ctx->have_result = false;
  return AC_OK;
}

extern "C" int ac_set_graph(ac_ctx* ctx, int which, int64_t nV, int64_t dV, const float* vf, int64_t nE, int64_t dE,
                            const float* ef, const int32_t* src, const int32_t* dst) {
  if (!ctx) {
    ac::set_global_error("ac_set_graph: null context");
    return AC_ERR_ARG;
  }
  if (which != AC_DATA && which != AC_PATTERN) return ctx->fail(AC_ERR_ARG, "ac_set_graph: which must be AC_DATA or AC_PATTERN");
  if (nV < 1 || nV > INT32_MAX || dV < 0 || nE < 0 || dE < 0)
    return ctx->fail(AC_ERR_ARG, "ac_set_graph: need 1 <= num_nodes < 2^31 and non-negative sizes");
  if ((nV * dV > 0 && !vf) || (nE * dE > 0 && !ef) || (nE > 0 && (!src || !dst)))
    return ctx->fail(AC_ERR_ARG, "ac_set_graph: null array for a non-empty graph");
  for (int64_t e = 0; e < nE; ++e)
    if (src[e] < 0 || src[e] >= nV || dst[e] < 0 || dst[e] >= nV)
      return ctx->fail(AC_ERR_ARG, "ac_set_graph: edge " + std::to_string(e) + " has an endpoint outside [0, " +
                                       std::to_string(nV) + ")");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (which == AC_DATA) return set_data(ctx, nV, dV, vf, nE, dE, ef, src, dst);
  return set_pattern(ctx, nV, dV, vf, nE, dE, ef, src, dst);
}

extern "C" int ac_set_graph64(ac_ctx* ctx, int which, const ac_graph64* g) {
  if (!ctx) {
    ac::set_global_error("ac_set_graph64: null context");
    return AC_ERR_ARG;
  }
  if (!g) return ctx->fail(AC_ERR_ARG, "ac_set_graph64: null graph");
  if (g->num_nodes < 1 || g->num_nodes > (uint64_t)INT32_MAX || g->num_edges > (uint64_t)INT32_MAX ||
      g->node_feat_dim > 4096 || g->edge_feat_dim > 4096)
    return ctx->fail(AC_ERR_ARG, "ac_set_graph64: sizes out of range (need 1 <= num_nodes < 2^31, edges < 2^31)");
  const uint64_t nV = g->num_nodes, nE = g->num_edges;
  std::vector<float> vf(nV * g->node_feat_dim), ef(nE * g->edge_feat_dim);
  std::vector<int32_t> s(nE), d(nE);
  ac::parallel_chunks((int64_t)vf.size(), [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i) vf[i] = (float)g->node_feats[i];
  });
  ac::parallel_chunks((int64_t)ef.size(), [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i) ef[i] = (float)g->edge_feats[i];
  });
  for (uint64_t e = 0; e < nE; ++e) {
    if (g->srcs[e] >= nV || g->dsts[e] >= nV)
      return ctx->fail(AC_ERR_ARG, "ac_set_graph64: edge " + std::to_string(e) + " has an endpoint >= num_nodes");
    s[e] = (int32_t)g->srcs[e];
    d[e] = (int32_t)g->dsts[e];
  }
  return ac_set_graph(ctx, which, (int64_t)nV, (int64_t)g->node_feat_dim, vf.data(), (int64_t)nE,
                      (int64_t)g->edge_feat_dim, ef.data(), s.data(), d.data());
}

static int walk_batch_of(const ac_ctx* ctx) { return ctx->walk_B > 0 ? ctx->walk_B : ac::walk_batch(); }

// Bundle lengths are padded to a multiple of the largest walk batch, so one layout serves every batch size.
constexpr int kMaxWalkBatch = 4;
// Bundle layouts kept per data graph (one per distinct G among recent patterns).  Patterns of up to 64 edges have
// at most 12 distinct G (G = 64 / (Qs / 4)), so a mix of such widths run one after another never rebuilds a layout.
constexpr size_t kBundleCache = 16;

// Selects the bundle layout for the current pattern width, building it if it is not cached: the
// in-edge runs, longest first (ties by first source, so that the bundles walked at the same time gather from nearby
// sources), are dealt out to bundles of G = 64 / (Qs / 4) runs, then the rows of every bundle are laid out on the
// device (k_build_bundles) and the CSR walk is re-targeted at them.
static int select_bundles(ac_ctx* ctx) {
  const int G = ac::walk_groups(ctx->Qs), Bt = kMaxWalkBatch;
  for (size_t i = 0; i < ctx->bsets.size(); ++i)
    if (ctx->bsets[i]->G == G) {
      std::unique_ptr<ac_ctx::BundleSet> hit = std::move(ctx->bsets[i]);
      ctx->bsets.erase(ctx->bsets.begin() + i);
      ctx->bsets.push_back(std::move(hit));
      ctx->bs = ctx->bsets.back().get();
      return AC_OK;
    }
  const std::vector<Run>& R = ctx->h_runs_in;
  const int64_t nr = (int64_t)R.size();
  std::vector<int> order(nr);
  for (int64_t i = 0; i < nr; ++i) order[i] = (int)i;
  const std::vector<int>& S0 = ctx->h_run_src0;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int la = R[a].ke - R[a].kb, lb = R[b].ke - R[b].kb;
    return la > lb || (la == lb && S0[a] < S0[b]);
  });
  const int64_t nb = (nr + G - 1) / G;
  std::vector<Bundle> bun(nb);
  std::vector<int> brun((size_t)std::max<int64_t>(1, nb * G), -1), gslot((size_t)std::max<int64_t>(1, nb * G), -1);
  int64_t base = 0;
  for (int64_t b = 0; b < nb; ++b) {
    int len = 0;
    for (int g = 0; g < G; ++g) {
      const int64_t i = b * G + g;
      if (i >= nr) break;
      const Run& r = R[order[i]];
      brun[i] = order[i];
      gslot[i] = r.slot;
      len = std::max(len, r.ke - r.kb);
    }
    len = (len + Bt - 1) / Bt * Bt;
    bun[b] = Bundle{base, len, 0};
    base += (int64_t)len * G;
  }
  if (base >= (int64_t)INT32_MAX) return ctx->fail(AC_ERR_ARG, "ac_run: the padded edge layout exceeds 2^31 rows");
  hipStream_t st = ctx->stream;
  ctx->bs = nullptr;
  if (ctx->bsets.size() >= kBundleCache) ctx->bsets.erase(ctx->bsets.begin());
  ctx->bsets.push_back(std::unique_ptr<ac_ctx::BundleSet>(new ac_ctx::BundleSet()));
  ac_ctx::BundleSet* S = ctx->bsets.back().get();
  S->G = G;
  S->nbundles = nb;
  S->brows = base;
  DevBuf<int> d_brun, rowmap;
  const size_t rows = (size_t)std::max<int64_t>(1, base), m = (size_t)std::max<int64_t>(1, ctx->de);
  int rc = AC_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == AC_OK)
      rc = ctx->fail(e == hipErrorOutOfMemory ? AC_ERR_OOM : AC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  chk(S->bundles.upload(bun.data(), bun.size(), st), "bundle upload");
  chk(d_brun.upload(brun.data(), brun.size(), st), "bundle upload");
  chk(S->gslot.upload(gslot.data(), gslot.size(), st), "bundle upload");
  chk(S->bsrc.alloc(rows), "bundle rows");
  chk(S->bdst.alloc(rows), "bundle rows");
  chk(S->beid.alloc(rows), "bundle rows");
  chk(rowmap.alloc(m), "bundle rows");
  chk(S->csr_kb.alloc(m), "bundle rows");
  if (rc != AC_OK) {
    ctx->bsets.pop_back();
    return rc;
  }
  ac::launch_build_bundles(nb, S->bundles.p, d_brun.p, ctx->runs_in.p, ctx->csc_sd.p, ctx->csc_eid.p, G, S->bsrc.p,
                           S->bdst.p, S->beid.p, rowmap.p, st);
  ac::launch_remap_ks(ctx->de, ctx->csr_ks.p, rowmap.p, S->csr_kb.p, st);
  chk(hipGetLastError(), "k_build_bundles");
  chk(hipStreamSynchronize(st), "k_build_bundles");  // d_brun / rowmap are freed on return
  if (rc != AC_OK) {
    ctx->bsets.pop_back();
    return rc;
  }
  ctx->bs = S;
  return AC_OK;
}

// Sizes every work buffer for the current (data, pattern) pair.
static int ensure_work(ac_ctx* ctx) {
  int rc = select_bundles(ctx);
  if (rc != AC_OK) return rc;
  const int64_t key[5] = {ctx->dn, ctx->de, ctx->P, ctx->Q, ctx->bs->brows};
  if (std::equal(key, key + 5, ctx->work_key)) return AC_OK;
  const int64_t n = ctx->dn;
  const int P = ctx->P, Q = ctx->Q, Qs = ctx->Qs;
  ctx->VT = ac::vertex_tile(P, Q);
  const int64_t ntiles = (std::max<int64_t>(ctx->n_act, 1) + ctx->VT - 1) / ctx->VT;
  {
    ac::VertexArgs va{};
    va.P = P;
    va.Q = Q;
    va.Qs = Qs;
    va.VT = ctx->VT;
    const int64_t all_tiles = (n + ctx->VT - 1) / ctx->VT;
    ctx->nb_vertex = vertex_blocks("AC_VERTEX_BLOCKS", ac::vertex_resident_blocks(ac::kModeA, va), 1, ntiles);
    ctx->nb_vertex_c = vertex_blocks("AC_VERTEX_BLOCKS_C", ac::vertex_resident_blocks(ac::kModeC, va), 1, ntiles);
    ctx->nb_vertex_o = vertex_blocks("AC_VERTEX_BLOCKS_O", ac::vertex_resident_blocks(ac::kModeOut, va), 2,
                                     all_tiles);
  }
  ctx->nb_col = (int)std::min<int64_t>((n + 255) / 256, 1024);
  ctx->nb_src = std::max(1, ac::walk_blocks(ctx->nruns_out, Qs));
  const int nb_l = std::max(ctx->nb_col, ctx->nb_src);
  HIPCHK(ctx, ctx->CE.alloc((size_t)std::max<int64_t>(1, ctx->bs->brows * Qs)));
  HIPCHK(ctx, ctx->CV.alloc((size_t)(n * P)));
  HIPCHK(ctx, ctx->FMaxN.alloc((size_t)(n * Qs)));
  HIPCHK(ctx, ctx->VFm.alloc((size_t)(n * Qs)));
  HIPCHK(ctx, ctx->minCEo.alloc((size_t)(n * Qs)));
  HIPCHK(ctx, ctx->lsO.alloc((size_t)(n * Qs)));
  ctx->VRs = gather_stride(Qs);
  HIPCHK(ctx, ctx->VR.alloc((size_t)(n * ctx->VRs)));
  HIPCHK(ctx, ctx->OUT.alloc((size_t)(n * P)));
  HIPCHK(ctx, ctx->carry_in.alloc((size_t)std::max(1, ctx->nslots_in) * Qs));
  HIPCHK(ctx, ctx->fix_part.alloc((size_t)std::max(1, ctx->nfix_parts) * Qs));
  HIPCHK(ctx, ctx->carry_out.alloc((size_t)std::max(1, ctx->nslots_out) * Qs * 2));
  HIPCHK(ctx, ctx->qpart.alloc((size_t)(ctx->nb_vertex + 1) * std::max(1, Q)));
  HIPCHK(ctx, ctx->ppart.alloc((size_t)(ctx->nb_vertex_c + 1) * P));
  HIPCHK(ctx, ctx->lpart.alloc((size_t)nb_l * std::max(P, std::max(Q, 1))));
  HIPCHK(ctx, ctx->qstat.alloc((size_t)3 * std::max(1, Q)));
  HIPCHK(ctx, ctx->pstat.alloc((size_t)P));
  HIPCHK(ctx, ctx->lse_cv.alloc((size_t)P));
  HIPCHK(ctx, ctx->lse_ce.alloc((size_t)std::max(1, Q)));
  HIPCHK(ctx, ctx->iso.alloc((size_t)P));
  HIPCHK(ctx, ctx->apv.alloc((size_t)ctx->nb_col * P));
  HIPCHK(ctx, ctx->api.alloc((size_t)ctx->nb_col * P));
  // Padding columns of the row-major [n][Qs] arrays are never read as data; keep them zero.
  HIPCHK(ctx, hipMemsetAsync(ctx->VR.p, 0, ctx->VR.n * sizeof(float), ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->qstat.p, 0, ctx->qstat.n * sizeof(QStat), ctx->stream));
  std::copy(key, key + 5, ctx->work_key);
  return AC_OK;
}

// Records the end of `stage` when profiling (a no-op otherwise).
static int mark(ac_ctx* ctx, int stage) {
  if (!ctx->profiling) return AC_OK;
  if (ctx->pev_used == ctx->pev.size()) {
    hipEvent_t e;
    HIPCHK(ctx, hipEventCreate(&e));
    ctx->pev.push_back(e);
    ctx->pev_stage.push_back(0);
  }
  ctx->pev_stage[ctx->pev_used] = stage;
  HIPCHK(ctx, hipEventRecord(ctx->pev[ctx->pev_used++], ctx->stream));
  return AC_OK;
}

#define MARK(stage)                  \
  do {                               \
    int _rc = mark(ctx, (stage));    \
    if (_rc != AC_OK) return _rc;    \
  } while (0)

// Checks the context and sizes the work buffers for a run of the current (data, pattern) pair.
static int prepare_run(ac_ctx* ctx) {
  if (!ctx) {
    ac::set_global_error("ac_run: null context");
    return AC_ERR_ARG;
  }
  if (!ctx->has_data || !ctx->has_patt) return ctx->fail(AC_ERR_STATE, "ac_run: set both the data and the pattern graph first");
  if (ctx->pdV != ctx->dV || ctx->pdE != ctx->dE)
    return ctx->fail(AC_ERR_ARG, "ac_run: feature dimensions of the data and pattern graphs differ");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (ctx->cnull && ctx->tape_K > 0)
    return ctx->fail(AC_ERR_ARG, "ac_run: AC_OPT_CNULL is not differentiated by ac_backward; use it without AC_OPT_TAPE");
  int rc = ensure_work(ctx);
  if (rc != AC_OK || !ctx->cnull || ctx->Q == 0) return rc;
  // Cnull (test.py:93-95): normprob over each pattern's edges of max(||edge features||_2, max_e normprob_e(CE)).  The
  // second term is a log-probability (<= 0) and the norm is >= 0, so the maximum is the norm, exactly; cnull[q] is the
  // log-softmax of the pattern's edge feature norms.
  std::vector<double> c((size_t)ctx->Q, 0.0);
  const int dE = ctx->pdE;
  for (size_t gi = 0; gi + 1 < ctx->cnull_groups.size(); ++gi) {
    const int q0 = ctx->cnull_groups[gi], q1 = ctx->cnull_groups[gi + 1];
    if (q0 >= q1) continue;
    double m = -std::numeric_limits<double>::infinity();
    for (int q = q0; q < q1; ++q) {
      double ss = 0.0;
      for (int k = 0; k < dE; ++k) ss += (double)ctx->pefeat[(size_t)q * dE + k] * (double)ctx->pefeat[(size_t)q * dE + k];
      c[q] = std::sqrt(ss);
      m = std::max(m, c[q]);
    }
    double sum = 0.0;
    for (int q = q0; q < q1; ++q) sum += std::exp(c[q] - m);
    for (int q = q0; q < q1; ++q) c[q] = std::log(std::exp(c[q] - m) / sum);
  }
  HIPCHK(ctx, ctx->d_cnull.upload(c.data(), c.size(), ctx->stream));
  return AC_OK;
}

static const double* cnull_of(const ac_ctx* ctx) { return ctx->cnull && ctx->Q > 0 ? ctx->d_cnull.p : nullptr; }

// Copies the live iteration state (FMaxN, VFm, the given scalars, pstat) into checkpoint c.
static int checkpoint(ac_ctx* ctx, int c, const QStat* qs) {
  const size_t nq = (size_t)ctx->dn * ctx->Qs, Q = (size_t)std::max(1, ctx->Q), P = (size_t)ctx->P;
  hipStream_t st = ctx->stream;
  HIPCHK(ctx, hipMemcpyAsync(ctx->ckF.p + c * nq, ctx->FMaxN.p, nq * sizeof(float), hipMemcpyDeviceToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(ctx->ckV.p + c * nq, ctx->VFm.p, nq * sizeof(float), hipMemcpyDeviceToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(ctx->ckQ.p + c * Q, qs, Q * sizeof(QStat), hipMemcpyDeviceToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(ctx->ckP.p + c * P, ctx->pstat.p, P * sizeof(Lse2), hipMemcpyDeviceToDevice, st));
  return AC_OK;
}

// Buffers of the adjoint tape for a run of n_iter iterations: checkpoints of iterations 0, K, 2K, ... and the
// argmax records of the prep.
static int ensure_tape(ac_ctx* ctx, int n_iter) {
  const int K = ctx->tape_K;
  // With K = 1 one more slot: the state after the last iteration (the run writes every state into its slot).
  const size_t nck = (size_t)std::max(1, (n_iter + K - 1) / K) + (K == 1 ? 1 : 0);
  const size_t nq = (size_t)ctx->dn * ctx->Qs, Q = (size_t)std::max(1, ctx->Q), P = (size_t)ctx->P;
  HIPCHK(ctx, ctx->FA0.alloc(nq));
  HIPCHK(ctx, ctx->AO.alloc(nq));
  HIPCHK(ctx, ctx->am0.alloc(P));
  HIPCHK(ctx, ctx->carg_in.alloc((size_t)std::max(1, ctx->nslots_in) * ctx->Qs));
  HIPCHK(ctx, ctx->ckF.alloc(nck * nq));
  HIPCHK(ctx, ctx->ckV.alloc(nck * nq));
  HIPCHK(ctx, ctx->ckQ.alloc(nck * Q));
  HIPCHK(ctx, ctx->ckP.alloc(nck * P));
  HIPCHK(ctx, ctx->apv.alloc((size_t)ctx->nb_col * P));
  HIPCHK(ctx, ctx->api.alloc((size_t)ctx->nb_col * P));
  if (K == 1) HIPCHK(ctx, ctx->tapeFA.alloc((size_t)std::max(1, n_iter) * nq));
  return AC_OK;
}

// Combines the split vertices' pieces of a plain (non-recording) walk.
static void launch_fixup32(ac_ctx* ctx, hipStream_t st) {
  ac::launch_walk_max_fixup32(ctx->nfix_chunks, ctx->fix_chunks.p, ctx->nfix_multi, ctx->fix_multi.p, ctx->carry_in.p,
                              ctx->Qs, ctx->fix_part.p, ctx->FMaxN.p, st);
}

// Prep (src/main.cu:107-148), between events ev[0] and ev[1].
static int enqueue_prep(ac_ctx* ctx) {
  Range range("prep");
  hipStream_t st = ctx->stream;
  const int64_t n = ctx->dn;
  const int P = ctx->P, Q = ctx->Q, Qs = ctx->Qs;
  QStat* qs_prep = ctx->qstat.p;

  HIPCHK(ctx, hipEventRecord(ctx->ev[0], st));
  MARK(-1);
  // ---- prep (src/main.cu:107-148)
  // Vertex and edge distances run at the same time, so each has its own pattern operands.
  const size_t vsd = ac::cdist_scratch_doubles(P, (int)ctx->dV), esd = ac::cdist_scratch_doubles(Qs, (int)ctx->dE);
  HIPCHK(ctx, ctx->cd_scratch.alloc(std::max<size_t>(1, vsd + esd)));
  double* cds_v = ctx->use_mfma_cdist ? ctx->cd_scratch.p : nullptr;
  double* cds_e = ctx->use_mfma_cdist ? ctx->cd_scratch.p + vsd : nullptr;
  if (Q > 0) {
    HIPCHK(ctx, ctx->bmin.alloc((size_t)std::max<int64_t>(1, ctx->bs->nbundles) * Qs));
    if (ctx->pack_costs) HIPCHK(ctx, ctx->bpack.alloc((size_t)std::max<int64_t>(1, ctx->bs->nbundles)));
    if (ctx->pack_costs) HIPCHK(ctx, ctx->CEp.alloc((size_t)std::max<int64_t>(1, ctx->bs->brows * (Qs / 4) * 3)));
  }
  // The vertex side (distances, column statistics, isolated columns) and the constant rows of the per-vertex arrays
  // go to the second stream, beside the edge distances on the first; the first waits for them (xev[1]) before the
  // per-source statistics, which reuse the partial buffer, and the prep scalars, which read lse_cv.
  hipStream_t ax = ctx->aux;
  HIPCHK(ctx, hipEventRecord(ctx->xev[0], st));
  HIPCHK(ctx, hipStreamWaitEvent(ax, ctx->xev[0], 0));
  ac::launch_cdist(n, ctx->d_perm.p, ctx->vfeat.p, (int)ctx->dV, ctx->pf_v.p, P, round8(P), P, ctx->CV.p, P,
                   ctx->has_wv ? ctx->wv.p : nullptr, cds_v, ax);
  ac::launch_col_lse_neg(n, P, ctx->CV.p, ctx->lpart.p, ctx->nb_col, ax);
  ac::launch_reduce_lse(ctx->nb_col, P, ctx->lpart.p, ctx->lse_cv.p, ax);
  if (ctx->n_iso > 0)
    ac::launch_iso_cols(ctx->n_iso, P, ctx->CV.p + ctx->n_act * P, (int)ctx->n_act, ctx->lpart.p, ctx->apv.p,
                        ctx->api.p, std::min(ctx->nb_col, 256), ctx->iso.p, ax);
  HIPCHK(ctx, hipMemcpyAsync(ctx->pstat.p, ctx->lse_cv.p, sizeof(Lse2) * P, hipMemcpyDeviceToDevice, ax));
  if (Q > 0) {
    // Only the constant rows: the per-source statistics write every row of a vertex with out-edges, and the prep
    // walk every row of a vertex with in-edges (internal order: both, in-edges only, out-edges only, isolated).
    const int64_t in_lo = ctx->n_both, in_hi = ctx->n_both + ctx->n_inonly, iso = ctx->n_act;
    for (const auto& r : {std::make_pair(in_lo, in_hi), std::make_pair(iso, n)}) {
      ac::launch_fill(ctx->minCEo.p + r.first * Qs, (r.second - r.first) * Qs, INFINITY, ax);
      ac::launch_fill(ctx->lsO.p + r.first * Qs, (r.second - r.first) * Qs, 0.f, ax);
      ac::launch_fill(ctx->VFm.p + r.first * Qs, (r.second - r.first) * Qs, -INFINITY, ax);
    }
    ac::launch_fill(ctx->FMaxN.p + in_hi * Qs, (n - in_hi) * Qs, -INFINITY, ax);
  }
  HIPCHK(ctx, hipEventRecord(ctx->xev[1], ax));
  if (Q > 0) {
    ac::launch_cdist(ctx->bs->brows, ctx->bs->beid.p, ctx->efeat.p, (int)ctx->dE, ctx->pf_e.p, Q, round8(Q), Qs, ctx->CE.p,
                     Qs, ctx->has_we ? ctx->we.p : nullptr, cds_e, st);
    // The bundle minima (a stream over CE) on the second stream, beside the per-source statistics (a gather over
    // CE, bound by its own instruction chains); the first stream waits for them (xev[3]) before the prep walk.
    HIPCHK(ctx, hipEventRecord(ctx->xev[2], st));
    HIPCHK(ctx, hipStreamWaitEvent(ax, ctx->xev[2], 0));
    ac::launch_bundle_min(ctx->bs->nbundles, ctx->bs->bundles.p, ctx->CE.p, Q, Qs, ctx->bmin.p,
                          ctx->pack_costs ? ctx->bpack.p : nullptr, ax);
    HIPCHK(ctx, hipEventRecord(ctx->xev[3], ax));
    HIPCHK(ctx, hipStreamWaitEvent(st, ctx->xev[1], 0));
    if (ctx->nruns_out > 0) {
      ac::launch_walk_src(ctx->nruns_out, ctx->runs_out.p, ctx->bs->csr_kb.p, ctx->CE.p, Q, Qs, ctx->minCEo.p,
                          ctx->lsO.p, ctx->VFm.p, ctx->carry_out.p, ctx->lpart.p, ctx->nb_src, st);
      ac::launch_walk_src_fixup(ctx->nsplit_out, ctx->splits_out.p, ctx->carry_out.p, Q, Qs, ctx->minCEo.p,
                                ctx->lsO.p, ctx->VFm.p, st);
      ac::launch_reduce_lse(ctx->nb_src, Q, ctx->lpart.p, ctx->lse_ce.p, st);
    } else {
      std::vector<Lse2> empty(Q, Lse2{-std::numeric_limits<double>::infinity(), 0.0});
      HIPCHK(ctx, hipMemcpyAsync(ctx->lse_ce.p, empty.data(), sizeof(Lse2) * Q, hipMemcpyHostToDevice, st));
      HIPCHK(ctx, hipStreamSynchronize(st));
    }
    ac::launch_prep_scalars(P, Q, ctx->d_ps.p, ctx->d_pd.p, ctx->lse_cv.p, ctx->lse_ce.p, qs_prep, st);
    HIPCHK(ctx, hipStreamWaitEvent(st, ctx->xev[3], 0));
    ctx->costs_packed = false;
    // The prep's max over in-edges of -CE reads every cost row once, so it also writes the 24-bit codes of the
    // bundles that pack (one pass over CE instead of a separate packing pass).
    const bool tape = ctx->tape_K > 0;
    ac::launch_walk_max(ctx->bs->nbundles, ctx->bs->bundles.p, ctx->bs->gslot.p, ctx->bs->bsrc.p, ctx->bs->bdst.p,
                        ctx->CE.p, nullptr, ctx->VRs, Q, Qs, qs_prep, ctx->FMaxN.p, ctx->carry_in.p,
                        walk_batch_of(ctx), st, tape ? ctx->FA0.p : nullptr, tape ? ctx->carg_in.p : nullptr,
                        nullptr, nullptr, ctx->pack_costs ? ctx->bpack.p : nullptr,
                        ctx->pack_costs ? ctx->CEp.p : nullptr);
    ctx->costs_packed = ctx->pack_costs;
    if (tape)
      ac::launch_walk_max_fixup(ctx->nsplit_in, ctx->splits_in.p, ctx->carry_in.p, Q, Qs, qs_prep, ctx->FMaxN.p, st,
                                ctx->carg_in.p, ctx->FA0.p);
    else
      launch_fixup32(ctx, st);
    if (tape) {
      // Rows of the minimum-cost out-edge of every (v, q) (the RMax argmax), for the adjoint.
      ac::launch_fill_int(ctx->AO.p, n * Qs, INT32_MAX, st);
      ac::launch_argmin_out(ctx->de, Q, Qs, ctx->bs->csr_kb.p, ctx->CE.p, ctx->minCEo.p, ctx->AO.p, st);
    }
  }
  HIPCHK(ctx, hipStreamWaitEvent(st, ctx->xev[1], 0));  // (again, for Q = 0) everything of the second stream
  if (ctx->tape_K > 0) {
    ac::launch_col_argmin(n, P, ctx->CV.p, ctx->apv.p, ctx->api.p, std::min(ctx->nb_col, 128), ctx->am0.p, st);
    int rc = checkpoint(ctx, 0, qs_prep);
    if (rc != AC_OK) return rc;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], st));
  MARK(AC_STAGE_PREP);
  return AC_OK;
}


static ac::VertexArgs vertex_args(ac_ctx* ctx) {
  ac::VertexArgs va{};
  va.n = ctx->n_act;
  va.perm = ctx->d_perm.p;
  va.f_end = ctx->n_both + ctx->n_inonly;
  va.o_lo = ctx->n_both;
  va.o_hi = ctx->n_both + ctx->n_inonly;
  va.o_end = ctx->n_act;
  va.iso = ctx->n_iso > 0 ? ctx->iso.p : nullptr;
  va.track_arg = ctx->tape_K > 0;
  va.P = ctx->P;
  va.Q = ctx->Q;
  va.Qs = ctx->Qs;
  va.VT = ctx->VT;
  va.CV = ctx->CV.p;
  va.FMaxN = ctx->FMaxN.p;
  va.VFm = ctx->VFm.p;
  va.VFm_out = ctx->VFm.p;
  va.minCEo = ctx->minCEo.p;
  va.lsO = ctx->lsO.p;
  va.VR = ctx->VR.p;
  va.VRs = ctx->VRs;
  va.out = ctx->OUT.p;
  va.ps = ctx->d_ps.p;
  va.pd = ctx->d_pd.p;
  va.inc_off = ctx->d_inc_off.p;
  va.inc = ctx->d_inc.p;
  va.pst = ctx->pstat.p;
  va.qpart = ctx->qpart.p;
  va.ppart = ctx->ppart.p;
  return va;
}

// One iteration (src/main.cu:160-245) reading the scalars of state `qs_state`, writing those of the new state into
// the parity slot of `it`.  Kernels only (no copies or fills), so it can be captured into a graph.  FA (nullable):
// the walk also records its argmax rows there (the split vertices' pieces are then combined with theirs); the maxima
// are the same as a plain walk's, bit for bit.
static int enqueue_iteration(ac_ctx* ctx, ac::VertexArgs& va, int it, const QStat*& qs_state, int* FA = nullptr) {
  hipStream_t st = ctx->stream;
  const int P = ctx->P, Q = ctx->Q, Qs = ctx->Qs;
  QStat* qs_new = ctx->qstat.p + (1 + (it & 1)) * std::max(1, Q);
  Range loop("loop");
  if (Q > 0) {
    va.first = it == 0;
    va.qs = qs_state;
    {
      Range r("step1+step3+step4a");  // VF, VR; VFmax, RE's normaliser, RMax; VRmax, FE's normaliser
      ac::launch_vertex(ac::kModeA, va, ctx->nb_vertex, st);
      MARK(AC_STAGE_VERTEX_A);
      ac::launch_reduce_q(ctx->nb_vertex + 1, Q, ctx->qpart.p, qs_new, st, cnull_of(ctx));
      MARK(AC_STAGE_REDUCE_A);
    }
    Range r("step2+step4b+step4c");  // FE messages, normalised, scatter-max by destination
    ac::launch_walk_max(ctx->bs->nbundles, ctx->bs->bundles.p, ctx->bs->gslot.p, ctx->bs->bsrc.p, ctx->bs->bdst.p,
                        ctx->CE.p, ctx->VR.p, ctx->VRs, Q, Qs, qs_new, ctx->FMaxN.p, ctx->carry_in.p,
                        walk_batch_of(ctx), st, FA, FA ? ctx->carg_in.p : nullptr, ctx->prune ? ctx->bmin.p : nullptr,
                        ctx->costs_packed ? ctx->CEp.p : nullptr, ctx->costs_packed ? ctx->bpack.p : nullptr);
    MARK(AC_STAGE_EDGE_WALK);
    if (FA)
      ac::launch_walk_max_fixup(ctx->nsplit_in, ctx->splits_in.p, ctx->carry_in.p, Q, Qs, qs_new, ctx->FMaxN.p, st,
                                ctx->carg_in.p, FA);
    else
      launch_fixup32(ctx, st);
    MARK(AC_STAGE_WALK_FIXUP);
    qs_state = qs_new;
  }
  va.first = 0;
  va.qs = qs_state;
  Range r5("step5");  // MU sums and MU's normaliser
  ac::launch_vertex(ac::kModeC, va, ctx->nb_vertex_c, st);
  MARK(AC_STAGE_VERTEX_C);
  ac::launch_reduce_lse(ctx->nb_vertex_c + 1, P, ctx->ppart.p, ctx->pstat.p, st);
  MARK(AC_STAGE_REDUCE_C);
  return AC_OK;
}

// Output of the current state (src/main.cu:247): MU[v][p] written directly in dn x pn order into OUT.
static int enqueue_output(ac_ctx* ctx, ac::VertexArgs& va, int done, const QStat* qs_state) {
  ac::VertexArgs vo = va;
  vo.n = ctx->dn;  // every row, isolated ones included, written to the caller's row
  vo.first = done == 0;
  vo.qs = qs_state;
  ac::launch_vertex(ac::kModeOut, vo, ctx->nb_vertex_o, ctx->stream);
  return AC_OK;
}

// A copy of the output columns [c0, c0 + ncol) taken after `t` iterations into dst[dn][ncol].
struct Snapshot {
  int t, c0, ncol;
  float* dst;
};

// Captured iteration pairs of one loop: one per parity of the pair's first iteration, because iteration `it`
// reads the scalar slot that iteration it - 1 wrote (slot 1 + ((it - 1) & 1)) and writes slot 1 + (it & 1), and a
// captured launch keeps the slot addresses it was captured with.  Each graph is kept with its executable until the
// loop's work has finished.
struct PairGraphs {
  hipGraph_t g[2] = {nullptr, nullptr};
  hipGraphExec_t x[2] = {nullptr, nullptr};
  void release() {
    for (int i = 0; i < 2; ++i) {
      if (x[i]) (void)hipGraphExecDestroy(x[i]);
      if (g[i]) (void)hipGraphDestroy(g[i]);
      x[i] = nullptr;
      g[i] = nullptr;
    }
  }
};

// A profiling tool loaded into the process (rocprofv3 names its tool library in ROCP_TOOL_LIBRARIES): graph replay
// is then left off, so that every iteration is an ordinary launch the tool traces.
static bool profiler_attached() {
  const char* t = std::getenv("ROCP_TOOL_LIBRARIES");
  return t && *t;
}

// The loop (src/main.cu:160-245) from the prep state: n_iter iterations, the snapshots (sorted by t) taken as their
// iteration counts are reached, and, without snapshots, the final output into OUT.  With `graph`, iterations
// 1, 2, ... are replayed in pairs from captured graphs (one per parity of the pair's first iteration) wherever no
// snapshot falls inside the pair.  Ends with event ev[2].
static int enqueue_loop(ac_ctx* ctx, int n_iter, const std::vector<Snapshot>& snaps, bool graph) {
  hipStream_t st = ctx->stream;
  ac::VertexArgs va = vertex_args(ctx);
  const QStat* qs_state = ctx->qstat.p;  // scalars finishing FMaxN / VFm as they currently are
  const int qstride = std::max(1, ctx->Q);
  size_t si = 0;
  auto take = [&](int t) -> int {
    while (si < snaps.size() && snaps[si].t == t) {
      const Snapshot& S = snaps[si++];
      int rc = enqueue_output(ctx, va, t, qs_state);
      if (rc != AC_OK) return rc;
      ac::launch_copy_cols(ctx->dn, ctx->P, S.c0, S.ncol, ctx->OUT.p, S.dst, st);
    }
    return AC_OK;
  };
  int rc = take(0);
  if (rc != AC_OK) return rc;
  if (graph && profiler_attached()) graph = false;
  PairGraphs pg;
  bool replayed = false;
  int it = 0;
  while (it < n_iter) {
    const bool pair = graph && it >= 1 && it + 1 < n_iter && !(si < snaps.size() && snaps[si].t == it + 1);
    if (pair) {
      const int par = it & 1;
      if (!pg.x[par]) {
        // The pair reads the slot iteration it - 1 wrote; it is where qs_state points now.
        HIPCHK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const QStat* q0 = qs_state;
        int rc1 = enqueue_iteration(ctx, va, it, q0);
        if (rc1 == AC_OK) rc1 = enqueue_iteration(ctx, va, it + 1, q0);
        const hipError_t e = hipStreamEndCapture(st, &pg.g[par]);
        if (rc1 == AC_OK && e != hipSuccess)
          rc1 = ctx->fail(AC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        if (rc1 == AC_OK) {
          const hipError_t ei = hipGraphInstantiate(&pg.x[par], pg.g[par], nullptr, nullptr, 0);
          if (ei != hipSuccess) {
            pg.x[par] = nullptr;
            rc1 = ctx->fail(AC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
          }
        }
        if (rc1 != AC_OK) {
          if (replayed) (void)hipStreamSynchronize(st);
          pg.release();
          return rc1;
        }
      }
      const hipError_t el = hipGraphLaunch(pg.x[par], st);
      if (el != hipSuccess) {
        (void)hipStreamSynchronize(st);
        pg.release();
        HIPCHK(ctx, el);
      }
      replayed = true;
      // The pair's second iteration, it + 1, wrote slot 1 + ((it + 1) & 1).
      qs_state = ctx->Q == 0 ? ctx->qstat.p : ctx->qstat.p + (1 + ((it + 1) & 1)) * qstride;
      it += 2;
    } else {
      rc = enqueue_iteration(ctx, va, it, qs_state);
      if (rc != AC_OK) break;
      ++it;
      if (ctx->tape_K > 0 && it % ctx->tape_K == 0 && it < n_iter) {
        rc = checkpoint(ctx, it / ctx->tape_K, qs_state);
        if (rc != AC_OK) break;
      }
    }
    rc = take(it);
    if (rc != AC_OK) break;
  }
  if (pg.x[0] || pg.x[1]) {
    // Executables and graphs are destroyed only after every replay has finished.
    const hipError_t es = hipStreamSynchronize(st);
    pg.release();
    HIPCHK(ctx, es);
  }
  if (rc != AC_OK) return rc;
  if (snaps.empty()) {
    rc = enqueue_output(ctx, va, n_iter, qs_state);
    if (rc != AC_OK) return rc;
  }
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], st));
  MARK(AC_STAGE_OUTPUT);
  return AC_OK;
}

static int enqueue_taped_loop(ac_ctx* ctx, int n_iter);

static int enqueue_run(ac_ctx* ctx, int n_iter) {
  int rc = prepare_run(ctx);
  if (rc != AC_OK) return rc;
  if (n_iter < 0) n_iter = ctx->P;  // src/main.cu:160
  ctx->tape_valid = false;
  if (ctx->tape_K > 0) {
    rc = ensure_tape(ctx, n_iter);
    if (rc != AC_OK) return rc;
  }
  ctx->pev_used = 0;
  rc = enqueue_prep(ctx);
  if (rc == AC_OK)
    rc = ctx->tape_K == 1 ? enqueue_taped_loop(ctx, n_iter)
                          : enqueue_loop(ctx, n_iter, {}, ctx->use_graph && !ctx->profiling && ctx->tape_K == 0);
  if (rc != AC_OK) return rc;
  ctx->have_result = true;
  ctx->last_iters = n_iter;
  ctx->tape_valid = ctx->tape_K > 0;
  ctx->tape_iters = n_iter;
  return AC_OK;
}

extern "C" int ac_run_async(ac_ctx* ctx, int n_iter) { return enqueue_run(ctx, n_iter); }

extern "C" int ac_sync(ac_ctx* ctx) {
  if (!ctx) return AC_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return AC_OK;
}

extern "C" int ac_run(ac_ctx* ctx, int n_iter) {
  int rc = enqueue_run(ctx, n_iter);
  if (rc != AC_OK) return rc;
  return ac_sync(ctx);
}

extern "C" int ac_get_mu(ac_ctx* ctx, float* host_out) {
  if (!ctx || !host_out) return ctx ? ctx->fail(AC_ERR_ARG, "ac_get_mu: null output") : AC_ERR_ARG;
  if (!ctx->have_result) return ctx->fail(AC_ERR_STATE, "ac_get_mu: no result; call ac_run first");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipMemcpyAsync(host_out, ctx->OUT.p, sizeof(float) * ctx->dn * ctx->P, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return AC_OK;
}

extern "C" int ac_get_match(ac_ctx* ctx, float log_tau, uint8_t* host_out) {
  if (!ctx || !host_out) return ctx ? ctx->fail(AC_ERR_ARG, "ac_get_match: null output") : AC_ERR_ARG;
  const int64_t n = ctx->dn;
  const int P = ctx->P;
  std::vector<float> mu((size_t)(n * P));
  int rc = ac_get_mu(ctx, mu.data());
  if (rc != AC_OK) return rc;
  std::vector<int64_t> arg(P, 0);
  for (int64_t v = 1; v < n; ++v)
    for (int p = 0; p < P; ++p)
      if (mu[v * P + p] > mu[arg[p] * P + p]) arg[p] = v;
  std::memset(host_out, 0, (size_t)(n * P));
  for (int p = 0; p < P; ++p) host_out[arg[p] * P + p] = mu[arg[p] * P + p] >= log_tau ? 2 : 1;
  return AC_OK;
}

extern "C" int ac_get_topk(ac_ctx* ctx, int k, int32_t* idx_out, float* val_out) {
  if (!ctx || !idx_out) return ctx ? ctx->fail(AC_ERR_ARG, "ac_get_topk: null output") : AC_ERR_ARG;
  if (k < 1 || (int64_t)k > ctx->dn)
    return ctx->fail(AC_ERR_ARG, "ac_get_topk: need 1 <= k <= number of data vertices");
  const int64_t n = ctx->dn;
  const int P = ctx->P;
  std::vector<float> mu((size_t)(n * P));
  int rc = ac_get_mu(ctx, mu.data());
  if (rc != AC_OK) return rc;
  // Per pattern vertex: the k best (value descending, vertex ascending on ties), by a bounded selection on the
  // calling thread (dn * pn comparisons, no worker threads).
  for (int64_t p = 0; p < P; ++p) {
    auto better = [&](int a, int b) {
      const float x = mu[(size_t)a * P + p], y = mu[(size_t)b * P + p];
      return x > y || (x == y && a < b);
    };
    std::vector<int> best;
    best.reserve((size_t)k + 1);
    for (int64_t v = 0; v < n; ++v) {
      if ((int)best.size() == k && !better((int)v, best.back())) continue;
      best.insert(std::upper_bound(best.begin(), best.end(), (int)v, better), (int)v);
      if ((int)best.size() > k) best.pop_back();
    }
    for (int r = 0; r < k; ++r) {
      idx_out[p * k + r] = best[r];
      if (val_out) val_out[p * k + r] = mu[(size_t)best[r] * P + p];
    }
  }
  return AC_OK;
}

extern "C" int ac_get_sizes(const ac_ctx* ctx, int64_t* out4) {
  if (!ctx || !out4) return AC_ERR_ARG;
  out4[0] = ctx->dn;
  out4[1] = ctx->de;
  out4[2] = ctx->P;
  out4[3] = ctx->Q;
  return AC_OK;
}

extern "C" int ac_get_cost_packing(ac_ctx* ctx, int64_t* out2) {
  if (!ctx || !out2) return AC_ERR_ARG;
  if (!ctx->have_result || !ctx->bs || ctx->Q == 0)
    return ctx->fail(AC_ERR_STATE, "ac_get_cost_packing: no run with pattern edges yet");
  const int64_t nb = ctx->bs->nbundles;
  int64_t packed = 0;
  if (ctx->costs_packed && nb > 0) {
    std::vector<unsigned> h((size_t)nb);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(h.data(), ctx->bpack.p, sizeof(unsigned) * (size_t)nb, hipMemcpyDeviceToHost));
    for (unsigned b : h) packed += b != 0u;
  }
  out2[0] = nb;
  out2[1] = packed;
  return AC_OK;
}

extern "C" int ac_get_timing(const ac_ctx* ctx, double* prep_ms, double* loop_ms) {
  if (!ctx) return AC_ERR_ARG;
  if (!ctx->have_result) return AC_ERR_STATE;
  float a = 0.f, b = 0.f;
  if (hipEventSynchronize(ctx->ev[2]) != hipSuccess || hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]) != hipSuccess ||
      hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]) != hipSuccess)
    return AC_ERR_HIP;
  if (prep_ms) *prep_ms = a;
  if (loop_ms) *loop_ms = b;
  return AC_OK;
}

extern "C" int ac_run_profiled(ac_ctx* ctx, int n_iter, double* stage_ms) {
  if (!ctx || !stage_ms) return ctx ? ctx->fail(AC_ERR_ARG, "ac_run_profiled: null stage_ms") : AC_ERR_ARG;
  ctx->profiling = true;
  int rc = enqueue_run(ctx, n_iter);
  ctx->profiling = false;
  if (rc != AC_OK) return rc;
  rc = ac_sync(ctx);
  if (rc != AC_OK) return rc;
  for (int s = 0; s < AC_NSTAGES; ++s) stage_ms[s] = 0.0;
  for (size_t i = 1; i < ctx->pev_used; ++i) {
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->pev[i - 1], ctx->pev[i]));
    const int s = ctx->pev_stage[i];
    if (s >= 0 && s < AC_NSTAGES) stage_ms[s] += ms;
  }
  return AC_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Weights.

extern "C" int ac_set_weights(ac_ctx* ctx, const float* wv, const float* we) {
  if (!ctx) {
    ac::set_global_error("ac_set_weights: null context");
    return AC_ERR_ARG;
  }
  if (!ctx->has_data) return ctx->fail(AC_ERR_STATE, "ac_set_weights: set the data graph first (it fixes the widths)");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto set = [&](const float* w, int64_t dim, bool& has, DevBuf<double>& d, std::vector<double>& h) -> int {
    if (!w) {
      has = false;
      h.clear();
      return AC_OK;
    }
    h.assign(w, w + dim);
    for (double x : h)
      if (!(x >= 0.0) || !std::isfinite(x)) return ctx->fail(AC_ERR_ARG, "ac_set_weights: weights must be finite and >= 0");
    HIPCHK(ctx, d.upload(h.data(), h.size(), ctx->stream));
    has = true;
    return AC_OK;
  };
  int rc = set(wv, ctx->dV, ctx->has_wv, ctx->wv, ctx->h_wv);
  if (rc == AC_OK) rc = set(we, ctx->dE, ctx->has_we, ctx->we, ctx->h_we);
  ctx->have_result = false;
  return rc;
}

// ---------------------------------------------------------------------------------------------------------------
// Options.

extern "C" int ac_set_option(ac_ctx* ctx, int option, int value) {
  if (!ctx) return AC_ERR_ARG;
  switch (option) {
    case AC_OPT_WALK_BATCH:
      if (value != 0 && value != 1 && value != 2 && value != 4)
        return ctx->fail(AC_ERR_ARG, "ac_set_option: the walk batch is 1, 2 or 4 (0: default)");
      ctx->walk_B = value;
      return AC_OK;
    case AC_OPT_GRAPH:
      ctx->use_graph = value != 0;
      return AC_OK;
    case AC_OPT_SLAB_EDGES:
      if (value < 4 || value > 256) return ctx->fail(AC_ERR_ARG, "ac_set_option: slab edges must be in [4, 256]");
      ctx->slab_edges = value;
      return AC_OK;
    case AC_OPT_TAPE:
      if (value < 0 || value > 1024) return ctx->fail(AC_ERR_ARG, "ac_set_option: checkpoint interval in [0, 1024]");
      ctx->tape_K = value;
      ctx->tape_valid = false;
      return AC_OK;
    case AC_OPT_PRUNE:
      ctx->prune = value != 0;
      return AC_OK;
    case AC_OPT_CDIST_MFMA:
      ctx->use_mfma_cdist = value != 0;
      return AC_OK;
    case AC_OPT_PACK_COSTS:
      ctx->pack_costs = value != 0;
      return AC_OK;
    case AC_OPT_CNULL:
      ctx->cnull = value != 0;
      ctx->have_result = false;
      return AC_OK;
    default:
      return ctx->fail(AC_ERR_ARG, "ac_set_option: unknown option " + std::to_string(option));
  }
}

extern "C" int ac_get_option(const ac_ctx* ctx, int option, int* value) {
  if (!ctx || !value) return AC_ERR_ARG;
  switch (option) {
    case AC_OPT_WALK_BATCH:
      *value = walk_batch_of(ctx);
      return AC_OK;
    case AC_OPT_GRAPH:
      *value = ctx->use_graph ? 1 : 0;
      return AC_OK;
    case AC_OPT_SLAB_EDGES:
      *value = ctx->slab_edges;
      return AC_OK;
    case AC_OPT_TAPE:
      *value = ctx->tape_K;
      return AC_OK;
    case AC_OPT_PRUNE:
      *value = ctx->prune ? 1 : 0;
      return AC_OK;
    case AC_OPT_CDIST_MFMA:
      *value = ctx->use_mfma_cdist ? 1 : 0;
      return AC_OK;
    case AC_OPT_PACK_COSTS:
      *value = ctx->pack_costs ? 1 : 0;
      return AC_OK;
    case AC_OPT_CNULL:
      *value = ctx->cnull ? 1 : 0;
      return AC_OK;
    default:
      return AC_ERR_ARG;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Batched scoring.

extern "C" int ac_batch_set_patterns(ac_ctx* ctx, int k, const int32_t* v_off, const int32_t* e_off, int64_t dV,
                                     const float* vfeat, int64_t dE, const float* efeat, const int32_t* src,
                                     const int32_t* dst) {
  if (!ctx) {
    ac::set_global_error("ac_batch_set_patterns: null context");
    return AC_ERR_ARG;
  }
  if (k < 1 || !v_off || !e_off || dV < 0 || dE < 0)
    return ctx->fail(AC_ERR_ARG, "ac_batch_set_patterns: need k >= 1, offsets and non-negative widths");
  if (v_off[0] != 0 || e_off[0] != 0) return ctx->fail(AC_ERR_ARG, "ac_batch_set_patterns: offsets start at 0");
  std::vector<ac_ctx::BatchPattern> bp((size_t)k);
  for (int i = 0; i < k; ++i) {
    const int64_t P = (int64_t)v_off[i + 1] - v_off[i], Q = (int64_t)e_off[i + 1] - e_off[i];
    if (P < 1 || P > 256 || Q < 0 || Q > 256)
      return ctx->fail(AC_ERR_ARG, "ac_batch_set_patterns: pattern " + std::to_string(i) +
                                       " needs 1..256 vertices and at most 256 edges");
    bp[i] = ac_ctx::BatchPattern{(int)P, (int)Q, v_off[i], e_off[i], 0};
  }
  const int64_t nv = v_off[k], ne = e_off[k];
  if ((nv * dV > 0 && !vfeat) || (ne * dE > 0 && !efeat) || (ne > 0 && (!src || !dst)))
    return ctx->fail(AC_ERR_ARG, "ac_batch_set_patterns: null array for non-empty patterns");
  for (int i = 0; i < k; ++i)
    for (int64_t e = bp[i].e0; e < bp[i].e0 + bp[i].Q; ++e)
      if (src[e] < 0 || src[e] >= bp[i].P || dst[e] < 0 || dst[e] >= bp[i].P)
        return ctx->fail(AC_ERR_ARG, "ac_batch_set_patterns: edge " + std::to_string(e - bp[i].e0) + " of pattern " +
                                         std::to_string(i) + " has an endpoint outside the pattern");
  ctx->bp = std::move(bp);
  ctx->b_vfeat.assign(vfeat, vfeat + nv * dV);
  ctx->b_efeat.assign(efeat, efeat + ne * dE);
  ctx->b_src.assign(src, src + ne);
  ctx->b_dst.assign(dst, dst + ne);
  ctx->b_dV = dV;
  ctx->b_dE = dE;
  ctx->have_batch = false;
  return AC_OK;
}

// Slabs: patterns in descending order of iterations (ties by index), each slab filled while its edges stay within
// `cap` and its vertices within 256.
static std::vector<std::vector<int>> pack_slabs(const ac_ctx* ctx, const std::vector<int>& iters, int cap) {
  const int k = (int)ctx->bp.size();
  std::vector<int> order(k);
  for (int i = 0; i < k; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return iters[a] > iters[b]; });
  std::vector<std::vector<int>> slabs;
  int sq = 0, sp = 0;
  for (int i : order) {
    const auto& B = ctx->bp[i];
    if (slabs.empty() || sq + B.Q > cap || sp + B.P > 256) {
      slabs.emplace_back();
      sq = sp = 0;
    }
    slabs.back().push_back(i);
    sq += B.Q;
    sp += B.P;
  }
  return slabs;
}

// Makes the disjoint union of a slab's patterns the context's pattern; col0[j] receives member j's first column.
static int set_slab_pattern(ac_ctx* ctx, const std::vector<int>& members, std::vector<int>* col0) {
  const int64_t dV = ctx->b_dV, dE = ctx->b_dE;
  std::vector<float> vf, ef;
  std::vector<int32_t> s, d;
  int base = 0;
  if (col0) col0->clear();
  for (int i : members) {
    const auto& B = ctx->bp[i];
    if (col0) col0->push_back(base);
    vf.insert(vf.end(), ctx->b_vfeat.begin() + B.v0 * dV, ctx->b_vfeat.begin() + (B.v0 + B.P) * dV);
    ef.insert(ef.end(), ctx->b_efeat.begin() + B.e0 * dE, ctx->b_efeat.begin() + (B.e0 + B.Q) * dE);
    for (int64_t e = B.e0; e < B.e0 + B.Q; ++e) {
      s.push_back(base + ctx->b_src[e]);
      d.push_back(base + ctx->b_dst[e]);
    }
    base += B.P;
  }
  const int rc = ac_set_graph(ctx, AC_PATTERN, base, dV, vf.data(), (int64_t)s.size(), dE, ef.data(), s.data(),
                              d.data());
  if (rc != AC_OK) return rc;
  // Cnull is normalised over each member pattern's own edges.
  ctx->cnull_groups.assign(1, 0);
  for (int i : members) ctx->cnull_groups.push_back(ctx->cnull_groups.back() + ctx->bp[i].Q);
  return AC_OK;
}

static int batch_iters(ac_ctx* ctx, const int32_t* n_iter, std::vector<int>& iters) {
  iters.resize(ctx->bp.size());
  for (size_t i = 0; i < ctx->bp.size(); ++i) iters[i] = (n_iter && n_iter[i] >= 0) ? n_iter[i] : ctx->bp[i].P;
  return AC_OK;
}

// Batched runs and the autotuner run untaped: they never feed ac_backward, and the tape buffers are sized by
// ac_run for its own pattern, not for a slab.  The guard restores the caller's AC_OPT_TAPE.
struct TapeOff {
  ac_ctx* c;
  int saved;
  explicit TapeOff(ac_ctx* ctx) : c(ctx), saved(ctx->tape_K) {
    c->tape_K = 0;
    c->tape_valid = false;
  }
  ~TapeOff() { c->tape_K = saved; }
};

static int batch_check(ac_ctx* ctx, const char* who) {
  if (!ctx->has_data) return ctx->fail(AC_ERR_STATE, std::string(who) + ": set the data graph first");
  if (ctx->bp.empty()) return ctx->fail(AC_ERR_STATE, std::string(who) + ": set the patterns first");
  if (ctx->b_dV != ctx->dV || ctx->b_dE != ctx->dE)
    return ctx->fail(AC_ERR_ARG, std::string(who) + ": feature dimensions of the data graph and the patterns differ");
  return AC_OK;
}

extern "C" int ac_batch_run(ac_ctx* ctx, const int32_t* n_iter) {
  if (!ctx) return AC_ERR_ARG;
  int rc = batch_check(ctx, "ac_batch_run");
  if (rc != AC_OK) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TapeOff untaped(ctx);
  std::vector<int> iters;
  batch_iters(ctx, n_iter, iters);
  const std::vector<std::vector<int>> slabs = pack_slabs(ctx, iters, ctx->slab_edges);
  int64_t total = 0;
  for (auto& B : ctx->bp) {
    B.out_off = total;
    total += ctx->dn * B.P;
  }
  ctx->have_batch = false;
  HIPCHK(ctx, ctx->bout.alloc((size_t)total));
  double prep = 0.0, loop = 0.0;
  int64_t upd = 0, done = 0;
  for (const auto& members : slabs) {
    std::vector<int> col0;
    rc = set_slab_pattern(ctx, members, &col0);
    if (rc != AC_OK) return rc;
    rc = prepare_run(ctx);
    if (rc != AC_OK) return rc;
    std::vector<Snapshot> snaps;
    int n_max = 0, sq = 0;
    for (size_t j = 0; j < members.size(); ++j) {
      const auto& B = ctx->bp[members[j]];
      snaps.push_back(Snapshot{iters[members[j]], col0[j], B.P, ctx->bout.p + B.out_off});
      n_max = std::max(n_max, iters[members[j]]);
      sq += B.Q;
      upd += ctx->de * B.Q * iters[members[j]];
    }
    std::stable_sort(snaps.begin(), snaps.end(), [](const Snapshot& a, const Snapshot& b) { return a.t < b.t; });
    done += ctx->de * sq * n_max;
    ctx->pev_used = 0;
    rc = enqueue_prep(ctx);
    if (rc == AC_OK) rc = enqueue_loop(ctx, n_max, snaps, ctx->use_graph);
    if (rc != AC_OK) return rc;
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[2]));
    float a = 0.f, b = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
    HIPCHK(ctx, hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
    prep += a;
    loop += b;
  }
  ctx->have_result = false;  // OUT holds the last slab, not a pattern set with ac_set_graph
  ctx->b_info[0] = (int64_t)ctx->bp.size();
  ctx->b_info[1] = (int64_t)slabs.size();
  ctx->b_info[2] = upd;
  ctx->b_info[3] = done;
  ctx->b_prep_ms = prep;
  ctx->b_loop_ms = loop;
  ctx->have_batch = true;
  return AC_OK;
}

extern "C" int ac_batch_get_mu(ac_ctx* ctx, int i, float* host_out) {
  if (!ctx || !host_out) return ctx ? ctx->fail(AC_ERR_ARG, "ac_batch_get_mu: null output") : AC_ERR_ARG;
  if (!ctx->have_batch) return ctx->fail(AC_ERR_STATE, "ac_batch_get_mu: no result; call ac_batch_run first");
  if (i < 0 || i >= (int)ctx->bp.size()) return ctx->fail(AC_ERR_ARG, "ac_batch_get_mu: no pattern " + std::to_string(i));
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const auto& B = ctx->bp[i];
  HIPCHK(ctx, hipMemcpyAsync(host_out, ctx->bout.p + B.out_off, sizeof(float) * ctx->dn * B.P, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return AC_OK;
}

extern "C" int ac_batch_info(const ac_ctx* ctx, int64_t* out4) {
  if (!ctx || !out4) return AC_ERR_ARG;
  if (!ctx->have_batch) return AC_ERR_STATE;
  std::copy(ctx->b_info, ctx->b_info + 4, out4);
  return AC_OK;
}

extern "C" int ac_batch_timing(const ac_ctx* ctx, double* prep_ms, double* loop_ms) {
  if (!ctx) return AC_ERR_ARG;
  if (!ctx->have_batch) return AC_ERR_STATE;
  if (prep_ms) *prep_ms = ctx->b_prep_ms;
  if (loop_ms) *loop_ms = ctx->b_loop_ms;
  return AC_OK;
}

extern "C" int ac_batch_autotune(ac_ctx* ctx, int probe_iters, int32_t* choice_out, double* est_ms_out) {
  if (!ctx) return AC_ERR_ARG;
  int rc = batch_check(ctx, "ac_batch_autotune");
  if (rc != AC_OK) return rc;
  if (probe_iters < 1 || probe_iters > 64) return ctx->fail(AC_ERR_ARG, "ac_batch_autotune: probe_iters in [1, 64]");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TapeOff untaped(ctx);
  std::vector<int> iters;
  batch_iters(ctx, nullptr, iters);
  const int caps[3] = {64, 128, 256};
  const int batches[3] = {1, 2, 4};
  const int saved_B = ctx->walk_B;
  double best = std::numeric_limits<double>::infinity();
  int best_cap = ctx->slab_edges, best_B = walk_batch_of(ctx);
  hipStream_t st = ctx->stream;
  for (int cap : caps) {
    const std::vector<std::vector<int>> slabs = pack_slabs(ctx, iters, cap);
    if (cap > caps[0] && slabs.size() == pack_slabs(ctx, iters, cap / 2).size()) continue;  // same packing
    rc = set_slab_pattern(ctx, slabs[0], nullptr);
    if (rc == AC_OK) rc = prepare_run(ctx);
    if (rc == AC_OK) {
      ctx->pev_used = 0;
      rc = enqueue_prep(ctx);
    }
    if (rc != AC_OK) break;
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
    float prep_ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&prep_ms, ctx->ev[0], ctx->ev[1]));
    const int q0 = std::max(1, ctx->Q);
    for (int B : batches) {
      ctx->walk_B = B;
      ac::VertexArgs va = vertex_args(ctx);
      const QStat* qs_state = ctx->qstat.p;
      HIPCHK(ctx, hipEventRecord(ctx->ev[1], st));
      for (int it = 0; it < probe_iters && rc == AC_OK; ++it) rc = enqueue_iteration(ctx, va, it, qs_state);
      if (rc != AC_OK) break;
      HIPCHK(ctx, hipEventRecord(ctx->ev[2], st));
      HIPCHK(ctx, hipEventSynchronize(ctx->ev[2]));
      float ms = 0.f;
      HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
      // Per slab: prep in proportion to its edges, iterations in proportion to edges x iterations (probe rates).
      const double it_ms_per_q = ms / probe_iters / q0, prep_ms_per_q = prep_ms / q0;
      double est = 0.0;
      for (const auto& members : slabs) {
        int sq = 0, n_max = 0;
        for (int i : members) {
          sq += ctx->bp[i].Q;
          n_max = std::max(n_max, iters[i]);
        }
        est += prep_ms_per_q * std::max(1, sq) + it_ms_per_q * std::max(1, sq) * n_max;
      }
      if (est < best) {
        best = est;
        best_cap = cap;
        best_B = B;
      }
    }
    if (rc != AC_OK) break;
  }
  ctx->walk_B = saved_B;
  ctx->have_result = false;
  if (rc != AC_OK) return rc;
  ctx->slab_edges = best_cap;
  ctx->walk_B = best_B;
  if (choice_out) {
    choice_out[0] = best_cap;
    choice_out[1] = best_B;
  }
  if (est_ms_out) *est_ms_out = best;
  return AC_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Adjoint.

// One iteration of the recompute, from the recorded state In into state Out (separate buffers: nothing is copied),
// the walk recording its argmax rows into FA.  Out's rows that no kernel writes (FMaxN of vertices without
// in-edges, VFm of vertices without out-edges) must already hold -inf.
struct StateBufs {
  float* F;
  float* V;
  QStat* qs;
  Lse2* pst;
};

static int recompute_iteration(ac_ctx* ctx, int it, const StateBufs& In, const StateBufs& Out, int* FA) {
  hipStream_t st = ctx->stream;
  const int P = ctx->P, Q = ctx->Q, Qs = ctx->Qs;
  const size_t nq = (size_t)ctx->dn * Qs;
  ac::VertexArgs va = vertex_args(ctx);
  if (Q > 0) {
    va.first = it == 0;
    va.qs = In.qs;
    va.FMaxN = In.F;
    va.VFm = In.V;
    va.VFm_out = Out.V;
    va.pst = In.pst;
    ac::launch_vertex(ac::kModeA, va, ctx->nb_vertex, st);
    MARK(AC_STAGE_VERTEX_A);
    ac::launch_reduce_q(ctx->nb_vertex + 1, Q, ctx->qpart.p, Out.qs, st);
    MARK(AC_STAGE_REDUCE_A);
    ac::launch_walk_max(ctx->bs->nbundles, ctx->bs->bundles.p, ctx->bs->gslot.p, ctx->bs->bsrc.p, ctx->bs->bdst.p,
                        ctx->CE.p, ctx->VR.p, ctx->VRs, Q, Qs, Out.qs, Out.F, ctx->carry_in.p, walk_batch_of(ctx),
                        st, FA, ctx->carg_in.p, ctx->prune ? ctx->bmin.p : nullptr,
                        ctx->costs_packed ? ctx->CEp.p : nullptr, ctx->costs_packed ? ctx->bpack.p : nullptr);
    MARK(AC_STAGE_EDGE_WALK);
    ac::launch_walk_max_fixup(ctx->nsplit_in, ctx->splits_in.p, ctx->carry_in.p, Q, Qs, Out.qs, Out.F, st,
                              ctx->carg_in.p, FA);
    MARK(AC_STAGE_WALK_FIXUP);
  } else {
    HIPCHK(ctx, hipMemcpyAsync(Out.F, In.F, nq * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(Out.V, In.V, nq * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(Out.qs, In.qs, sizeof(QStat), hipMemcpyDeviceToDevice, st));
    MARK(AC_STAGE_VERTEX_A);
  }
  va.first = 0;
  va.qs = Out.qs;
  va.FMaxN = Out.F;
  va.VFm = Out.V;
  va.VFm_out = Out.V;
  ac::launch_vertex(ac::kModeC, va, ctx->nb_vertex_c, st);
  MARK(AC_STAGE_VERTEX_C);
  ac::launch_reduce_lse(ctx->nb_vertex_c + 1, P, ctx->ppart.p, Out.pst, st);
  MARK(AC_STAGE_REDUCE_C);
  return AC_OK;
}

// The constant rows of nslots consecutive iteration states (FMaxN and VFm slots of dn x Qs floats): -inf in the
// FMaxN rows of vertices without in-edges and in the VFm rows of vertices without out-edges, which no iteration
// writes (internal order: both, in-edges only, out-edges only, isolated), as the prep sets them in the live state.
static void fill_constant_rows(ac_ctx* ctx, float* F, float* V, int nslots, hipStream_t st) {
  const int64_t Qs = ctx->Qs, nq = ctx->dn * Qs;
  const int64_t in_lo = ctx->n_both, in_hi = ctx->n_both + ctx->n_inonly, iso = ctx->n_act, n = ctx->dn;
  ac::launch_fill_slots(F, V, nslots, nq, in_hi * Qs, (n - in_hi) * Qs, in_lo * Qs, (in_hi - in_lo) * Qs, iso * Qs,
                        (n - iso) * Qs, st);
}

// The iterations of a run with a checkpoint at every iteration (AC_OPT_TAPE = 1): iteration it reads the state in
// slot it and writes the next state into slot it + 1, its walk recording its argmax rows into tape slot it (the
// maxima are the same as a plain walk's, bit for bit), so no state is copied and ac_backward recomputes nothing.
// Slot 0 holds the prep state (enqueue_prep); the constant rows of the other slots (FMaxN without in-edges, VFm
// without out-edges) are -inf, written once here.  The output pass reads the last slot.
static int enqueue_taped_loop(ac_ctx* ctx, int n_iter) {
  Range loop("loop");
  hipStream_t st = ctx->stream;
  const size_t nq = (size_t)ctx->dn * ctx->Qs, Q1 = (size_t)std::max(1, ctx->Q), P = (size_t)ctx->P;
  auto slot = [&](int c) {
    return StateBufs{ctx->ckF.p + c * nq, ctx->ckV.p + c * nq, ctx->ckQ.p + c * Q1, ctx->ckP.p + c * P};
  };
  fill_constant_rows(ctx, ctx->ckF.p + nq, ctx->ckV.p + nq, n_iter, st);
  for (int it = 0; it < n_iter; ++it) {
    const int rc = recompute_iteration(ctx, it, slot(it), slot(it + 1), ctx->tapeFA.p + (size_t)it * nq);
    if (rc != AC_OK) return rc;
  }
  const StateBufs S = slot(n_iter);
  ac::VertexArgs va = vertex_args(ctx);
  va.FMaxN = S.F;
  va.VFm = S.V;
  va.VFm_out = S.V;
  va.pst = S.pst;
  const int rc = enqueue_output(ctx, va, n_iter, S.qs);
  if (rc != AC_OK) return rc;
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], st));
  MARK(AC_STAGE_OUTPUT);
  return AC_OK;
}

// MU (FP64, internal vertex order) of a recorded state into MUd, by the vertex pass's own evaluation.
static void launch_mu(ac_ctx* ctx, const ac::BwdState& S, hipStream_t st) {
  ac::VertexArgs va = vertex_args(ctx);
  va.n = ctx->dn;
  va.FMaxN = S.FMaxN;
  va.VFm = const_cast<float*>(S.VFm);  // read only in this mode
  va.qs = S.qs;
  va.pst = S.pst;
  va.first = S.first;
  va.mu_out = ctx->MUd.p;
  ac::launch_vertex(ac::kModeMU, va, ctx->nb_vertex_o, st);
}

// Host side of the weight gradient from the outer-product sums A, B (K x dim) and C (K) and the pattern features
// pf (K x dim): dL/dw_k = sum_q A[q][k] - 2 pf[q][k] B[q][k] + pf[q][k]^2 C[q].
static void combine_outer(const std::vector<double>& s, const std::vector<float>& pf, int K, int dim, float* out) {
  for (int k = 0; k < dim; ++k) {
    double g = 0.0;
    for (int q = 0; q < K; ++q) {
      const double p = pf[(size_t)q * dim + k];
      g += s[(size_t)q * dim + k] - 2.0 * p * s[(size_t)K * dim + (size_t)q * dim + k] + p * p * s[2 * (size_t)K * dim + q];
    }
    out[k] = (float)g;
  }
}

extern "C" int ac_backward(ac_ctx* ctx, const float* grad_mu, float* grad_wv, float* grad_we) {
  if (!ctx) {
    ac::set_global_error("ac_backward: null context");
    return AC_ERR_ARG;
  }
  if (!grad_mu) return ctx->fail(AC_ERR_ARG, "ac_backward: null grad_mu");
  if (!ctx->tape_valid || !ctx->have_result)
    return ctx->fail(AC_ERR_STATE, "ac_backward: no taped run; set AC_OPT_TAPE > 0 and call ac_run first");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t n = ctx->dn;
  const int P = ctx->P, Q = ctx->Q, Qs = ctx->Qs, K = ctx->tape_K, n_iter = ctx->tape_iters;
  const int Q1 = std::max(1, Q);
  const size_t nq = (size_t)n * Qs, np = (size_t)n * P;
  const int nparts = (int)std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, 2048));
  const int64_t brows = ctx->bs->brows;
  // grad_mu is in the caller's vertex order; the adjoint works in the internal one.
  HIPCHK(ctx, ctx->gMUin.upload(grad_mu, np, st));
  HIPCHK(ctx, ctx->gMU.alloc(np));
  ac::launch_gather_rows(ctx->dn, ctx->P, ctx->d_perm.p, ctx->gMUin.p, ctx->gMU.p, st);
  HIPCHK(ctx, ctx->gCV.alloc(np));
  for (DevBuf<float>* b : {&ctx->gEF, &ctx->gER, &ctx->gVR, &ctx->gVF}) HIPCHK(ctx, b->alloc(nq));
  HIPCHK(ctx, ctx->gCE.alloc((size_t)std::max<int64_t>(1, brows) * Qs));
  HIPCHK(ctx, ctx->gVRfix.alloc(nq));
  HIPCHK(ctx, ctx->gmax.alloc((size_t)n_iter + 1));
  HIPCHK(ctx, ctx->gERsum.alloc(nq));
  HIPCHK(ctx, ctx->cD.alloc(nq));
  HIPCHK(ctx, ctx->MUd.alloc(np));
  HIPCHK(ctx, ctx->colsum.alloc((size_t)P));
  HIPCHK(ctx, ctx->bpart.alloc((size_t)nparts * std::max(4 * Q1, P)));
  HIPCHK(ctx, ctx->red4.alloc((size_t)4 * Q1));
  HIPCHK(ctx, ctx->cq.alloc((size_t)Q1));
  // With a checkpoint at every iteration the run recorded every walk's argmax rows, and nothing is recomputed: no
  // segment buffers.
  const bool tape_full = K == 1;
  if (!tape_full) {
    HIPCHK(ctx, ctx->sgF.alloc((size_t)(K + 1) * nq));
    HIPCHK(ctx, ctx->sgV.alloc((size_t)(K + 1) * nq));
    HIPCHK(ctx, ctx->sgQ.alloc((size_t)(K + 1) * Q1));
    HIPCHK(ctx, ctx->sgP.alloc((size_t)(K + 1) * P));
    HIPCHK(ctx, ctx->tapeFA.alloc((size_t)K * nq));
  }
  HIPCHK(ctx, hipMemsetAsync(ctx->gCV.p, 0, np * sizeof(float), st));
  // gVF / gVR of the state after the last one: zero (k_bwd_route of the first step back reads their negations).
  HIPCHK(ctx, hipMemsetAsync(ctx->gVF.p, 0, nq * sizeof(float), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->gVR.p, 0, nq * sizeof(float), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->gCE.p, 0, ctx->gCE.n * sizeof(float), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->cD.p, 0, nq * sizeof(double), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->gERsum.p, 0, nq * sizeof(float), st));
  // The fixed-point sums start at zero; the vertex kernel of each step clears the ones it adds.
  HIPCHK(ctx, hipMemsetAsync(ctx->gVRfix.p, 0, nq * sizeof(unsigned long long), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->gmax.p, 0, ((size_t)n_iter + 1) * sizeof(unsigned), st));

  int fix_shift = 0;
  while (((int64_t)1 << (23 + fix_shift)) < ctx->max_outdeg) ++fix_shift;
  ac::BwdGraph G{n, P, Q, Qs, ctx->CV.p, ctx->minCEo.p, ctx->lsO.p, ctx->d_ps.p, ctx->d_pd.p, ctx->d_inc_off.p,
                 ctx->d_inc.p, fix_shift};
  HIPCHK(ctx, ctx->vpart.alloc((size_t)ac::bwd_vtx_blocks(G) * std::max(1, P)));
  // Segment states: state 0 is the checkpoint itself, states 1..K live in the segment buffers, whose constant rows
  // (-inf) are set once here and never written afterwards.
  if (!tape_full) fill_constant_rows(ctx, ctx->sgF.p, ctx->sgV.p, K + 1, st);
  const int nck = n_iter == 0 ? 0 : (n_iter + K - 1) / K;
  bool have_mu = false;  // MUd holds MU of the output state of the next iteration to reverse
  for (int c = nck - 1; c >= 0; --c) {
    const int i0 = c * K, i1 = std::min(i0 + K, n_iter);
    auto checkpoint_state = [&](int cc) {
      return StateBufs{ctx->ckF.p + cc * nq, ctx->ckV.p + cc * nq, ctx->ckQ.p + (size_t)cc * Q1,
                       ctx->ckP.p + (size_t)cc * P};
    };
    auto state = [&](int s) {
      if (s == 0) return checkpoint_state(c);
      if (tape_full) return checkpoint_state(c + 1);  // the run wrote every state into its slot
      return StateBufs{ctx->sgF.p + s * nq, ctx->sgV.p + s * nq, ctx->sgQ.p + (size_t)s * Q1, ctx->sgP.p + (size_t)s * P};
    };
    // Argmax rows of iteration i's walk: recorded by the run (full tape) or by the recompute below.
    auto tape_of = [&](int i) { return ctx->tapeFA.p + (size_t)(tape_full ? i : i - i0) * nq; };
    // Recompute the segment from its checkpoint, keeping every state and the walk's argmax rows.
    if (!tape_full)
      for (int i = i0; i < i1; ++i) {
        const int s = i - i0;
        const int rc = recompute_iteration(ctx, i, state(s), state(s + 1), tape_of(i));
        if (rc != AC_OK) return rc;
      }
    // Reverse through the segment.
    for (int i = i1 - 1; i >= i0; --i) {
      const size_t s = (size_t)(i - i0);
      const StateBufs B0 = state((int)s), B1 = state((int)s + 1);
      const ac::BwdState S0{B0.F, B0.V, B0.qs, B0.pst, i == 0};
      const ac::BwdState S1{B1.F, B1.V, B1.qs, B1.pst, 0};
      if (!have_mu) launch_mu(ctx, S1, st);
      // Column sums of gMU: from the previous step's vertex kernel, or (first step back) of the caller's gradient.
      if (have_mu)
        ac::launch_sum_parts(ac::bwd_vtx_blocks(G), P, ctx->vpart.p, ctx->colsum.p, st);
      else
        ac::launch_colsum(n, P, ctx->gMU.p, ctx->bpart.p, nparts, ctx->colsum.p, st);
      ac::launch_bwd_out(G, ctx->MUd.p, ctx->colsum.p, ctx->gMU.p, ctx->gCV.p, st);
      unsigned* gmax = ctx->gmax.p + i;
      ac::launch_bwd_route(G, S1, ctx->gMU.p, 1, ctx->gVF.p, ctx->gVR.p, ctx->gEF.p, ctx->gER.p, gmax, ctx->gERsum.p,
                           ctx->bpart.p, nparts, ctx->red4.p, st);
      launch_mu(ctx, S0, st);  // MU of the input state, reused by the next step back
      have_mu = true;
      if (Q == 0) HIPCHK(ctx, hipMemsetAsync(ctx->gVR.p, 0, nq * sizeof(float), st));
      if (Q == 0) HIPCHK(ctx, hipMemsetAsync(ctx->gVF.p, 0, nq * sizeof(float), st));
      const ac::FmaxTerms FE{ctx->gEF.p, tape_of(i), ctx->bs->bsrc.p, brows, ctx->gCE.p, ctx->gVRfix.p,
                             gmax};
      ac::launch_bwd_in(G, S0, ctx->MUd.p, S1.qs, ctx->red4.p, ctx->gER.p, ctx->gVR.p, ctx->gVF.p, ctx->cD.p, FE, st);
      ac::launch_bwd_vtx(G, ctx->gVR.p, ctx->gVF.p, ctx->gVRfix.p, gmax, ctx->gMU.p, ctx->vpart.p, st);
    }
  }
  // Prep (src/main.cu:107-148): FMax0 / RMax0 from -CE normalised by LSE0 = log sum_e exp(-CE), filled from
  // MU0 = LSM(-CV).
  const ac::BwdState S0{ctx->ckF.p, ctx->ckV.p, ctx->ckQ.p, ctx->ckP.p, 1};
  std::vector<double> cq(Q1, 0.0);
  if (Q > 0) {
    ac::launch_bwd_route(G, S0, nullptr, 0, ctx->gVF.p, ctx->gVR.p, ctx->gEF.p, ctx->gER.p, nullptr, ctx->gERsum.p,
                         ctx->bpart.p, nparts, ctx->red4.p, st);
    ac::launch_bwd_fscatter(G, ctx->gEF.p, ctx->FA0.p, ctx->bs->bsrc.p, brows, ctx->gCE.p, st);
    ac::launch_bwd_rterms(G, ctx->gERsum.p, ctx->AO.p, brows, ctx->gCE.p, st);
    ac::launch_bwd_prep_fill(G, ctx->red4.p, ctx->am0.p, ctx->gMU.p, st);
    std::vector<double> red4((size_t)4 * Q);
    std::vector<QStat> qs0(Q);
    HIPCHK(ctx, hipMemcpyAsync(red4.data(), ctx->red4.p, red4.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipMemcpyAsync(qs0.data(), ctx->ckQ.p, Q * sizeof(QStat), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    for (int q = 0; q < Q; ++q) {
      const double lse0 = qs0[q].mFE + qs0[q].lFE;
      cq[q] = std::isfinite(lse0) ? (red4[q] + red4[Q + q]) * std::exp(-lse0) : 0.0;
    }
  }
  HIPCHK(ctx, hipMemcpyAsync(ctx->cq.p, cq.data(), Q1 * sizeof(double), hipMemcpyHostToDevice, st));
  if (!have_mu) launch_mu(ctx, S0, st);
  ac::launch_colsum(n, P, ctx->gMU.p, ctx->bpart.p, nparts, ctx->colsum.p, st);
  ac::launch_bwd_out(G, ctx->MUd.p, ctx->colsum.p, ctx->gMU.p, ctx->gCV.p, st);

  // dL/dw: gC / (2 C) against the squared feature differences, as outer-product sums over rows.
  ac::launch_bwd_vertex_coef((int64_t)np, ctx->CV.p, ctx->gCV.p, st);
  const int dV = (int)ctx->dV, dE = (int)ctx->dE;
  const int onp = 512;
  std::vector<float> gwv(dV, 0.f), gwe(dE, 0.f);
  if (dV > 0) {
    const size_t w = (size_t)2 * P * dV + P;
    HIPCHK(ctx, ctx->opart.alloc((size_t)onp * w));
    HIPCHK(ctx, ctx->osum.alloc(w));
    ac::launch_outer_acc(n, ctx->d_perm.p, ctx->vfeat.p, dV, ctx->gCV.p, P, P, ctx->opart.p, onp, st);
    ac::launch_sum_parts(onp, (int)w, ctx->opart.p, ctx->osum.p, st);
    std::vector<double> sv(w);
    HIPCHK(ctx, hipMemcpyAsync(sv.data(), ctx->osum.p, w * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    combine_outer(sv, ctx->pvfeat, P, dV, gwv.data());
  }
  if (dE > 0 && Q > 0 && brows > 0) {
    ac::launch_bwd_edge_coef(brows, Q, Qs, ctx->bs->beid.p, ctx->bs->bsrc.p, ctx->CE.p, ctx->cD.p, ctx->cq.p,
                             ctx->gCE.p, st);
    const size_t w = (size_t)2 * Q * dE + Q;
    HIPCHK(ctx, ctx->opart.alloc((size_t)onp * w));
    HIPCHK(ctx, ctx->osum.alloc(w));
    ac::launch_outer_acc(brows, ctx->bs->beid.p, ctx->efeat.p, dE, ctx->gCE.p, Qs, Q, ctx->opart.p, onp, st);
    ac::launch_sum_parts(onp, (int)w, ctx->opart.p, ctx->osum.p, st);
    std::vector<double> se(w);
    HIPCHK(ctx, hipMemcpyAsync(se.data(), ctx->osum.p, w * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    combine_outer(se, ctx->pefeat, Q, dE, gwe.data());
  }
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(st));
  if (grad_wv) std::copy(gwv.begin(), gwv.end(), grad_wv);
  if (grad_we) std::copy(gwe.begin(), gwe.end(), grad_we);
  return AC_OK;
}
