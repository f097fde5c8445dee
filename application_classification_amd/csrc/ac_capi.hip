// C-ABI implementation: the matcher context, graph upload and the driver's prep + loop (src/main.cu:66-257).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/ac_mi355x.h"
#include "ac_host.hpp"
#include "ac_internal.hpp"

using ac::Lse;
using ac::Lse2;
using ac::QPart;
using ac::QStat;
using ac::Bundle;
using ac::Run;
using ac::Split;

namespace {

// Edges per run of an edge walk (a tuning knob, AC_RUN_EDGES).  Vertices whose edge list is longer are split into
// pieces of this length.
int run_edges() {
  if (const char* e = std::getenv("AC_RUN_EDGES")) {
    const int v = std::atoi(e);
    if (v >= 4 && v <= 4096) return v;
  }
  return 64;
}

int round8(int x) { return (x + 7) / 8 * 8; }

// Row stride of the gathered VR rows: whole 128-byte lines (or an exact fraction of one), so that a gather of one
// row never straddles two lines.
int gather_stride(int Qs) {
  if (Qs > 32) return (Qs + 31) / 32 * 32;
  int s = 4;
  while (s < Qs) s *= 2;
  return s;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    if (count == n && p) return hipSuccess;
    release();
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = count;
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

void build_runs(const std::vector<int64_t>& off, int64_t n, std::vector<Run>& runs, std::vector<Split>& splits,
                int& nslots) {
  const int kRunEdges = run_edges();
  runs.clear();
  splits.clear();
  nslots = 0;
  bool open = false;
  int64_t b = 0, len = 0;
  auto close = [&](int64_t e) {
    if (open) runs.push_back(Run{(int)b, (int)e, -1, 0});
    open = false;
  };
  for (int64_t v = 0; v < n; ++v) {
    const int64_t deg = off[v + 1] - off[v];
    if (deg == 0) continue;
    if (deg > kRunEdges) {
      close(off[v]);
      const int pieces = (int)((deg + kRunEdges - 1) / kRunEdges);
      splits.push_back(Split{(int)v, nslots, pieces, 0});
      for (int i = 0; i < pieces; ++i) {
        const int64_t kb = off[v] + (int64_t)i * kRunEdges;
        runs.push_back(Run{(int)kb, (int)std::min(off[v + 1], kb + kRunEdges), nslots++, 0});
      }
      continue;
    }
    if (open && len + deg > kRunEdges) close(off[v]);
    if (!open) {
      open = true;
      b = off[v];
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
  std::string err;

  // Data graph (src/main.cu:8-19 / 49-62) plus its CSC and CSR views.
  bool has_data = false;
  int64_t dn = 0, dV = 0, de = 0, dE = 0;
  DevBuf<float> vfeat, efeat;
  DevBuf<int2> csc_sd, csr_ks;
  DevBuf<int> csc_eid;
  DevBuf<Run> runs_in, runs_out;
  DevBuf<Split> splits_in, splits_out;
  std::vector<Run> h_runs_in;
  int64_t nruns_in = 0, nruns_out = 0;
  int nsplit_in = 0, nsplit_out = 0, nslots_in = 0, nslots_out = 0;

  // Pattern graph.
  bool has_patt = false;
  int P = 0, Q = 0, Qs = 0, pdV = 0, pdE = 0;
  std::vector<float> pvfeat, pefeat;
  std::vector<int> ps, pd;
  DevBuf<double> pf_v, pf_e;
  DevBuf<int> d_ps, d_pd, d_inc_off, d_inc;

  // Work buffers.
  DevBuf<float> CE, CV, FMaxN, VFm, minCEo, lsO, VR, OUT;
  // Bundle layout of the in-edge runs (depends on the pattern through Qs): rows of CE, their endpoints and edge ids,
  // the row of every CSC position, and the CSR walk's {row, src} pairs.
  DevBuf<Bundle> bundles;
  DevBuf<int> brun, gslot, bsrc, bdst, beid, rowmap;
  DevBuf<int2> csr_kb;
  int64_t nbundles = 0, brows = 0;
  DevBuf<double> carry_in, carry_out;
  DevBuf<QPart> qpart;
  DevBuf<Lse> ppart, lpart;
  DevBuf<QStat> qstat;  // [3][Q]: prep state, then by iteration parity
  DevBuf<Lse2> pstat, lse_cv, lse_ce;
  int nb_vertex = 0, nb_col = 0, nb_src = 0, VT = 0, VRs = 0;
  int64_t work_key[4] = {-1, -1, -1, -1};

  bool have_result = false;
  int last_iters = 0;

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
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess ||
      hipEventCreate(&c->ev[2]) != hipSuccess) {
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
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->pev) (void)hipEventDestroy(e);
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
  // CSC (by destination, stable) and CSR (by source, stable) views.
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
  std::vector<int> csc_eid(nE), csc_pos(nE);
  {
    std::vector<int64_t> fill(in_off.begin(), in_off.end() - 1);
    for (int64_t e = 0; e < nE; ++e) {
      const int64_t k = fill[dst[e]]++;
      csc_eid[k] = (int)e;
      csc_sd[k] = make_int2(src[e], dst[e]);
      csc_pos[e] = (int)k;
    }
  }
  {
    std::vector<int64_t> fill(out_off.begin(), out_off.end() - 1);
    for (int64_t e = 0; e < nE; ++e) csr_ks[fill[src[e]]++] = make_int2(csc_pos[e], src[e]);
  }
  std::vector<Run> rin, rout;
  std::vector<Split> sin, sout;
  int slin = 0, slout = 0;
  build_runs(in_off, nV, rin, sin, slin);
  build_runs(out_off, nV, rout, sout, slout);

  ctx->has_data = false;
  hipStream_t st = ctx->stream;
  HIPCHK(ctx, ctx->vfeat.upload(vf, (size_t)(nV * dV), st));
  HIPCHK(ctx, ctx->efeat.upload(ef, (size_t)(nE * dE), st));
  HIPCHK(ctx, ctx->csc_sd.upload(csc_sd.data(), csc_sd.size(), st));
  HIPCHK(ctx, ctx->csr_ks.upload(csr_ks.data(), csr_ks.size(), st));
  HIPCHK(ctx, ctx->csc_eid.upload(csc_eid.data(), csc_eid.size(), st));
  HIPCHK(ctx, ctx->runs_in.upload(rin.data(), rin.size(), st));
  HIPCHK(ctx, ctx->runs_out.upload(rout.data(), rout.size(), st));
  HIPCHK(ctx, ctx->splits_in.upload(sin.data(), sin.size(), st));
  HIPCHK(ctx, ctx->splits_out.upload(sout.data(), sout.size(), st));
  ctx->dn = nV;
  ctx->dV = dV;
  ctx->de = nE;
  ctx->dE = dE;
  ctx->nruns_in = (int64_t)rin.size();
  ctx->nruns_out = (int64_t)rout.size();
  ctx->nsplit_in = (int)sin.size();
  ctx->nsplit_out = (int)sout.size();
  ctx->nslots_in = slin;
  ctx->nslots_out = slout;
  ctx->h_runs_in = std::move(rin);
  ctx->work_key[0] = -1;  // the bundle layout follows the graph, not only its sizes
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

// Deals the in-edge runs, longest first (ties in CSC order), out to bundles of G = 64 / (Qs / 4) runs, then lays
// out the rows of every bundle on the device (k_build_bundles) and re-targets the CSR walk at them.
static int build_bundles(ac_ctx* ctx) {
  const int G = ac::walk_groups(ctx->Qs), Bt = ac::walk_batch();
  const std::vector<Run>& R = ctx->h_runs_in;
  const int64_t nr = (int64_t)R.size();
  std::vector<int> order(nr);
  for (int64_t i = 0; i < nr; ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return R[a].ke - R[a].kb > R[b].ke - R[b].kb; });
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
  ctx->nbundles = nb;
  ctx->brows = base;
  HIPCHK(ctx, ctx->bundles.upload(bun.data(), bun.size(), st));
  HIPCHK(ctx, ctx->brun.upload(brun.data(), brun.size(), st));
  HIPCHK(ctx, ctx->gslot.upload(gslot.data(), gslot.size(), st));
  const size_t rows = (size_t)std::max<int64_t>(1, base), m = (size_t)std::max<int64_t>(1, ctx->de);
  HIPCHK(ctx, ctx->bsrc.alloc(rows));
  HIPCHK(ctx, ctx->bdst.alloc(rows));
  HIPCHK(ctx, ctx->beid.alloc(rows));
  HIPCHK(ctx, ctx->rowmap.alloc(m));
  HIPCHK(ctx, ctx->csr_kb.alloc(m));
  ac::launch_build_bundles(nb, ctx->bundles.p, ctx->brun.p, ctx->runs_in.p, ctx->csc_sd.p, ctx->csc_eid.p, G,
                           ctx->bsrc.p, ctx->bdst.p, ctx->beid.p, ctx->rowmap.p, st);
  ac::launch_remap_ks(ctx->de, ctx->csr_ks.p, ctx->rowmap.p, ctx->csr_kb.p, st);
  HIPCHK(ctx, hipGetLastError());
  return AC_OK;
}

// Sizes every work buffer for the current (data, pattern) pair.
static int ensure_work(ac_ctx* ctx) {
  const int64_t key[4] = {ctx->dn, ctx->de, ctx->P, ctx->Q};
  if (std::equal(key, key + 4, ctx->work_key)) return AC_OK;
  const int64_t n = ctx->dn, m = ctx->de;
  const int P = ctx->P, Q = ctx->Q, Qs = ctx->Qs;
  ctx->VT = ac::vertex_tile(P, Q);
  const int64_t ntiles = (n + ctx->VT - 1) / ctx->VT;
  ctx->nb_vertex = (int)std::min<int64_t>(ntiles, 8192);
  ctx->nb_col = (int)std::min<int64_t>((n + 255) / 256, 1024);
  ctx->nb_src = std::max(1, ac::walk_blocks(ctx->nruns_out, Qs));
  const int nb_l = std::max(ctx->nb_col, ctx->nb_src);
  int rc = build_bundles(ctx);
  if (rc != AC_OK) return rc;
  HIPCHK(ctx, ctx->CE.alloc((size_t)std::max<int64_t>(1, ctx->brows * Qs)));
  HIPCHK(ctx, ctx->CV.alloc((size_t)(n * P)));
  HIPCHK(ctx, ctx->FMaxN.alloc((size_t)(n * Qs)));
  HIPCHK(ctx, ctx->VFm.alloc((size_t)(n * Qs)));
  HIPCHK(ctx, ctx->minCEo.alloc((size_t)(n * Qs)));
  HIPCHK(ctx, ctx->lsO.alloc((size_t)(n * Qs)));
  ctx->VRs = gather_stride(Qs);
  HIPCHK(ctx, ctx->VR.alloc((size_t)(n * ctx->VRs)));
  HIPCHK(ctx, ctx->OUT.alloc((size_t)(n * P)));
  HIPCHK(ctx, ctx->carry_in.alloc((size_t)std::max(1, ctx->nslots_in) * Qs));
  HIPCHK(ctx, ctx->carry_out.alloc((size_t)std::max(1, ctx->nslots_out) * Qs * 2));
  HIPCHK(ctx, ctx->qpart.alloc((size_t)ctx->nb_vertex * std::max(1, Q)));
  HIPCHK(ctx, ctx->ppart.alloc((size_t)ctx->nb_vertex * P));
  HIPCHK(ctx, ctx->lpart.alloc((size_t)nb_l * std::max(P, std::max(Q, 1))));
  HIPCHK(ctx, ctx->qstat.alloc((size_t)3 * std::max(1, Q)));
  HIPCHK(ctx, ctx->pstat.alloc((size_t)P));
  HIPCHK(ctx, ctx->lse_cv.alloc((size_t)P));
  HIPCHK(ctx, ctx->lse_ce.alloc((size_t)std::max(1, Q)));
  // Padding columns of the row-major [n][Qs] arrays are never read as data; keep them zero.
  HIPCHK(ctx, hipMemsetAsync(ctx->VR.p, 0, ctx->VR.n * sizeof(float), ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->qstat.p, 0, ctx->qstat.n * sizeof(QStat), ctx->stream));
  std::copy(key, key + 4, ctx->work_key);
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

static int enqueue_run(ac_ctx* ctx, int n_iter) {
  if (!ctx) {
    ac::set_global_error("ac_run: null context");
    return AC_ERR_ARG;
  }
  if (!ctx->has_data || !ctx->has_patt) return ctx->fail(AC_ERR_STATE, "ac_run: set both the data and the pattern graph first");
  if (ctx->pdV != ctx->dV || ctx->pdE != ctx->dE)
    return ctx->fail(AC_ERR_ARG, "ac_run: feature dimensions of the data and pattern graphs differ");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = ensure_work(ctx);
  if (rc != AC_OK) return rc;
  if (n_iter < 0) n_iter = ctx->P;  // src/main.cu:160
  hipStream_t st = ctx->stream;
  const int64_t n = ctx->dn, m = ctx->de;
  const int P = ctx->P, Q = ctx->Q, Qs = ctx->Qs;
  QStat* qs_prep = ctx->qstat.p;
  QStat* qs_par[2] = {ctx->qstat.p + std::max(1, Q), ctx->qstat.p + 2 * std::max(1, Q)};

  HIPCHK(ctx, hipEventRecord(ctx->ev[0], st));
  ctx->pev_used = 0;
  MARK(-1);
  // ---- prep (src/main.cu:107-148)
  ac::launch_cdist(n, nullptr, ctx->vfeat.p, (int)ctx->dV, ctx->pf_v.p, P, round8(P), P, ctx->CV.p, P, st);
  ac::launch_col_lse_neg(n, P, ctx->CV.p, ctx->lpart.p, ctx->nb_col, st);
  ac::launch_reduce_lse(ctx->nb_col, P, ctx->lpart.p, ctx->lse_cv.p, st);
  HIPCHK(ctx, hipMemcpyAsync(ctx->pstat.p, ctx->lse_cv.p, sizeof(Lse2) * P, hipMemcpyDeviceToDevice, st));
  if (Q > 0) {
    ac::launch_cdist(ctx->brows, ctx->beid.p, ctx->efeat.p, (int)ctx->dE, ctx->pf_e.p, Q, round8(Q), Qs, ctx->CE.p,
                     Qs, st);
    ac::launch_fill(ctx->minCEo.p, n * Qs, INFINITY, st);
    ac::launch_fill(ctx->lsO.p, n * Qs, 0.f, st);
    ac::launch_fill(ctx->VFm.p, n * Qs, -INFINITY, st);
    ac::launch_fill(ctx->FMaxN.p, n * Qs, -INFINITY, st);
    if (ctx->nruns_out > 0) {
      ac::launch_walk_src(ctx->nruns_out, ctx->runs_out.p, ctx->csr_kb.p, ctx->CE.p, Q, Qs, ctx->minCEo.p,
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
    ac::launch_walk_max(ctx->nbundles, ctx->bundles.p, ctx->gslot.p, ctx->bsrc.p, ctx->bdst.p, ctx->CE.p, nullptr,
                        ctx->VRs, Q, Qs, qs_prep, ctx->FMaxN.p, ctx->carry_in.p, st);
    ac::launch_walk_max_fixup(ctx->nsplit_in, ctx->splits_in.p, ctx->carry_in.p, Q, Qs, qs_prep, ctx->FMaxN.p, st);
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], st));
  MARK(AC_STAGE_PREP);

  // ---- loop (src/main.cu:160-245)
  ac::VertexArgs va{};
  va.n = n;
  va.P = P;
  va.Q = Q;
  va.Qs = Qs;
  va.VT = ctx->VT;
  va.CV = ctx->CV.p;
  va.FMaxN = ctx->FMaxN.p;
  va.VFm = ctx->VFm.p;
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
  const QStat* qs_state = qs_prep;  // scalars finishing FMaxN / VFm as they currently are
  for (int it = 0; it < n_iter; ++it) {
    QStat* qs_new = qs_par[it & 1];
    if (Q > 0) {
      va.first = it == 0;
      va.qs = qs_state;
      ac::launch_vertex(ac::kModeA, va, ctx->nb_vertex, st);
      MARK(AC_STAGE_VERTEX_A);
      ac::launch_reduce_q(ctx->nb_vertex, Q, ctx->qpart.p, qs_new, st);
      MARK(AC_STAGE_REDUCE_A);
      ac::launch_walk_max(ctx->nbundles, ctx->bundles.p, ctx->gslot.p, ctx->bsrc.p, ctx->bdst.p, ctx->CE.p,
                          ctx->VR.p, ctx->VRs, Q, Qs, qs_new, ctx->FMaxN.p, ctx->carry_in.p, st);
      MARK(AC_STAGE_EDGE_WALK);
      ac::launch_walk_max_fixup(ctx->nsplit_in, ctx->splits_in.p, ctx->carry_in.p, Q, Qs, qs_new, ctx->FMaxN.p, st);
      MARK(AC_STAGE_WALK_FIXUP);
      qs_state = qs_new;
    }
    va.first = 0;
    va.qs = qs_state;
    ac::launch_vertex(ac::kModeC, va, ctx->nb_vertex, st);
    MARK(AC_STAGE_VERTEX_C);
    ac::launch_reduce_lse(ctx->nb_vertex, P, ctx->ppart.p, ctx->pstat.p, st);
    MARK(AC_STAGE_REDUCE_C);
  }
  // ---- output (src/main.cu:247): MU[v][p] written directly in dn x pn order; timed with the loop, as the
  // reference's timer covers its final transpose.
  va.first = n_iter == 0;
  va.qs = qs_state;
  ac::launch_vertex(ac::kModeOut, va, ctx->nb_vertex, st);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], st));
  MARK(AC_STAGE_OUTPUT);
  ctx->have_result = true;
  ctx->last_iters = n_iter;
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

extern "C" int ac_get_sizes(const ac_ctx* ctx, int64_t* out4) {
  if (!ctx || !out4) return AC_ERR_ARG;
  out4[0] = ctx->dn;
  out4[1] = ctx->de;
  out4[2] = ctx->P;
  out4[3] = ctx->Q;
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
