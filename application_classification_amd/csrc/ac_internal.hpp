// Shared definitions of the matcher's device buffers, scalar records and kernel launchers.
//
// Layout in HBM (one data graph, one pattern; n data vertices, m data edges, P pattern vertices, Q pattern edges):
//   - per-vertex x pattern-edge arrays are row-major [n][Qs], Qs = Q rounded up to a multiple of 4, so a row is a
//     whole number of float4s (the edge walk gathers whole rows);
//   - per-vertex x pattern-vertex arrays are row-major [n][P];
//   - the edge-cost matrix CE is row-major [rows][Qs] in "bundle" order: the in-edge runs (CSC order) are dealt
//     out G = 64 / (Qs / 4) to a bundle, and step s of the bundle's run g is row base + s * G + g, so one step of
//     the wave that walks a bundle reads G adjacent rows.  Runs shorter than their bundle are padded with rows of
//     +inf cost; bsrc / bdst hold each row's edge endpoints (padding: src 0, dst repeating the run's last one);
//   - CSR order travels as int2 pairs {row of the edge in CE, src}.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ac {

// log-sum-exp accumulator: value = m + log1p(t).  m is the largest element seen, t the sum of exp(x - m) over all
// other elements, so a log-probability near 0 is formed as (x - m) - log1p(t) without cancellation.
struct Lse {
  double m, t;
};
// Finished log-sum-exp: value = m + l.
struct Lse2 {
  double m, l;
};

// Per pattern edge q: the fill values of the edge max-reductions and the two edge log-sum-exps of one iteration
// (src/main.cu:197-224).  fillF = max_v VR[q][v] (fills FMax), fillR = max_v VF[q][v] (fills RMax).
struct QStat {
  double fillF, fillR;
  double mRE, lRE;  // log sum_e exp(RE[q][e])
  double mFE, lFE;  // log sum_e exp(FE[q][e])
  int vF, vR;       // the data vertex at which fillF / fillR is attained (lowest on ties; -1 in the prep state)
  double cthr;      // an edge whose cost is >= cthr cannot raise FMax above fillF (+inf: no bound)
};
// Per-block partial of QStat.
struct QPart {
  double vfmax, vrmax;
  Lse re, fe;
  int avf, avr;  // vertex of vfmax / vrmax
};

// Isolated data vertices (no in- and no out-edge) hold constant rows (FMaxN = VFm = -inf), so their MU column is
// -CV + (a per-p constant).  Per pattern vertex p: the log-sum-exp of -CV over them as (m, t), and the minimum of CV
// with its lowest (internal) vertex, computed once per run.
struct IsoCol {
  double m, t;
  float cmin;
  int vmin;  // -1: no isolated vertex
};

// A contiguous run of edges in CSC or CSR order covering whole vertices, or one piece of a vertex whose edge list
// is longer than a run (slot >= 0 then names its carry slot).
struct Run {
  int kb, ke, slot, pad;
};
// A vertex split into `count` pieces with carry slots [first, first + count).
struct Split {
  int v, first, count, pad;
};
// G runs walked side by side by one wave: rows [base, base + len * G) of CE / bsrc / bdst; len is a multiple of
// the walk's batch size.
struct Bundle {
  int64_t base;
  int len, pad;
};

constexpr int kWalkBlock = 256;
// Flag on a bundle row's source id: the row starts the next vertex of its run (vertices in a run are consecutive).
constexpr int kNewVertex = (int)0x80000000u;

struct VertexArgs {
  int64_t n;
  int P, Q, Qs, VT;
  const float* CV;      // [n][P]  vertex costs
  const float* FMaxN;   // [n][Qs] normalised destination maxima (-inf: no in-edge)
  float* VFm;           // [n][Qs] VF - min out-edge cost (-inf: no out-edge)
  float* VFm_out;       // mode A writes the new VFm here (== VFm: in place)
  const float* minCEo;  // [n][Qs] min out-edge cost (+inf: none)
  const float* lsO;     // [n][Qs] log1p(sum over other out-edges of exp(min - cost))
  float* VR;            // [n][VRs]
  int VRs;              // row stride of VR: a power of two up to 32 floats, else a multiple of 32 (whole lines)
  float* out;           // [n][P]
  double* mu_out;       // mode MU: [n][P] in internal vertex order
  const int* perm;      // mode OUT: internal row -> caller's row of `out` (nullptr: identity)
  // Constant rows, not read by mode A: rows >= f_end have no in-edge (FMaxN = -inf); rows in [o_lo, o_hi) and
  // >= o_end have no out-edge (VFm = -inf, minCEo = +inf, lsO = 0).
  int64_t f_end, o_lo, o_hi, o_end;
  const IsoCol* iso;    // [P] isolated vertices' columns (nullptr: none); modes A and C write their share as the
                        // extra partial slot gridDim.x
  const int* ps;
  const int* pd;
  const int* inc_off;  // [P + 1]
  const int* inc;      // incidence of each pattern vertex: (q << 1) | (1 if RMax term else 0), ascending q
  const QStat* qs;     // scalars that finish FMax / RMax of the state being read
  const Lse2* pst;     // log-sum-exp of each MU column
  QPart* qpart;        // [gridDim + 1][Q]
  Lse* ppart;          // [gridDim + 1][P]
  int first;           // MU = LSM(-CV): the prep state, no FMax / RMax terms
  int track_arg;       // mode A: record the vertices of the column maxima (QStat vF / vR)
};

enum VertexMode { kModeA = 0, kModeC = 1, kModeOut = 2, kModeMU = 3 };  // kModeMU: MU in FP64 into mu_out

// Launchers (ac_kernels.hip).  All asynchronous on `st`.
// scratch (may be null: no matrix-core path): cdist_scratch_doubles(Kstore, dim) doubles.
void launch_cdist(int64_t nrows, const int* idx, const float* feats, int dim, const double* pat, int K, int Kp,
                  int Kstore, float* out, int out_stride, const double* w, double* scratch, hipStream_t st);
size_t cdist_scratch_doubles(int Kstore, int dim);
void launch_col_lse_neg(int64_t n, int P, const float* X, Lse* part, int nblocks, hipStream_t st);
void launch_reduce_lse(int nparts, int ncol, const Lse* part, Lse2* out, hipStream_t st);
// cnull (nullptr: off): per pattern edge, the PNNL Cnull term subtracted from both fills of the new state.
void launch_reduce_q(int nparts, int Q, const QPart* part, QStat* out, hipStream_t st, const double* cnull = nullptr);
// Per column p of the n rows X[n][P]: IsoCol{LSE of -X, min X and its lowest row + base}; partial buffers of
// nparts * P entries each.
void launch_iso_cols(int64_t n, int P, const float* X, int base, Lse* lpart, float* pv, int* pi, int nparts,
                     IsoCol* out, hipStream_t st);
// Per-bundle least costs bmin[nbundles][Qs]; with bpack != nullptr also each bundle's cost-code base (0 = not packed).
void launch_bundle_min(int64_t nbundles, const Bundle* bundles, const float* CE, int Q, int Qs, float* bmin,
                       unsigned* bpack, hipStream_t st);
void launch_prep_scalars(int P, int Q, const int* ps, const int* pd, const Lse2* cv, const Lse2* ce, QStat* qs,
                         hipStream_t st);
void launch_fill(float* p, int64_t count, float value, hipStream_t st);
void launch_fill_slots(float* A, float* B, int64_t nslots, int64_t stride, int64_t a0, int64_t al, int64_t b0,
                       int64_t bl, int64_t c0, int64_t cl, hipStream_t st);
void launch_walk_src(int64_t nruns, const Run* runs, const int2* ks, const float* CE, int Q, int Qs, float* minCEo,
                     float* lsO, float* VFm, double* carry, Lse* gpart, int nblocks, hipStream_t st);
void launch_walk_src_fixup(int nsplit, const Split* splits, const double* carry, int Q, int Qs, float* minCEo,
                           float* lsO, float* VFm, hipStream_t st);
void launch_build_bundles(int64_t nbundles, const Bundle* bundles, const int* brun, const Run* runs, const int2* sd,
                          const int* eid, int G, int* bsrc, int* bdst, int* beid, int* rowmap, hipStream_t st);
void launch_remap_ks(int64_t m, const int2* ks, const int* rowmap, int2* out, hipStream_t st);
// CEp / bpack (launch_bundle_min): the walk reads the packed rows of the bundles whose base is nonzero.  CEp_out /
// bpack (the prep walk, no pruning): the walk reads the FP32 rows and writes those bundles' packed rows to CEp_out.
void launch_walk_max(int64_t nbundles, const Bundle* bundles, const int* gslot, const int* bsrc, const int* bdst,
                     const float* CE, const float* VR, int VRs, int Q, int Qs, const QStat* qs, float* FMaxN,
                     double* carry, int B, hipStream_t st, int* FA = nullptr, int* carg = nullptr,
                     const float* bmin = nullptr, const unsigned* CEp = nullptr, const unsigned* bpack = nullptr,
                     unsigned* CEp_out = nullptr);
// Plain-path fixup (FP32 normalised pieces): chunks of at most kFixChunk pieces, then the vertices of several
// chunks from their partials (part: one [Qs] row per partial).
constexpr int kFixChunk = 16;
void launch_walk_max_fixup32(int nchunk, const Split* chunks, int nms, const Split* ms, const double* carry, int Qs,
                             float* part, float* FMaxN, hipStream_t st);
void launch_walk_max_fixup(int nsplit, const Split* splits, const double* carry, int Q, int Qs, const QStat* qs,
                           float* FMaxN, hipStream_t st, const int* carg = nullptr, int* FA = nullptr);
void launch_vertex(int mode, const VertexArgs& a, int nblocks, hipStream_t st);
// Blocks of vertex pass `mode` resident at once on the current device (0 if unknown); a.VT/P/Q/Qs set its LDS.
int vertex_resident_blocks(int mode, const VertexArgs& a);
// out[i][c] = in[perm[i]][c] for the n rows of ncol floats.
void launch_gather_rows(int64_t n, int ncol, const int* perm, const float* in, float* out, hipStream_t st);
void launch_copy_cols(int64_t n, int P, int c0, int ncol, const float* src, float* dst, hipStream_t st);
int walk_blocks(int64_t nruns, int Qs);
int walk_groups(int Qs);  // runs per bundle: 64 / (Qs / 4)
int walk_batch();         // default steps per batch of the bundle walk: 1, 2 or 4 (bundle lengths are multiples of it)
int vertex_tile(int P, int Q);

}  // namespace ac
