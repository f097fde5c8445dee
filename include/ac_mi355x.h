/*
 * ac_mi355x.h — C ABI of the MI355X ApplicationClassification matcher.
 *
 * The matcher scores a data graph against a pattern graph and produces, for every pattern vertex p and data
 * vertex v, the log-probability MU[v][p] that v plays the role of p.  Every entry point here replaces one piece of
 * the reference's own interface; each declaration cites the reference file:line it replaces.  There is no other
 * FFI in the reference: its only caller of the ac:: operators is its driver (src/main.cu:66-257).
 *
 * Conventions
 *  - Every function returns an int status: AC_OK (0) or one of the AC_ERR_* codes.  No exception crosses the ABI.
 *    ac_last_error(ctx) (or ac_global_error() for calls without a context) returns a NUL-terminated message.
 *  - Values are FP32, vertex and edge ids are int32, sizes are int64.  The reference used FP64/uint64
 *    (src/helpers.hxx:3-4); the FP32 path is checked against an FP64 restatement at rtol 1e-5 / atol 1e-8.
 *  - "stream" arguments are hipStream_t passed as void* (NULL = default stream).  Device pointers are plain
 *    device addresses (hipMalloc, or a torch.cuda tensor's data_ptr()).
 *  - One host thread per context; contexts are independent.
 */
#ifndef AC_MI355X_H
#define AC_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AC_OK 0
#define AC_ERR_ARG 1   /* bad argument: null pointer, size mismatch, id out of range, ...          */
#define AC_ERR_OOM 2   /* device or host allocation failed                                        */
#define AC_ERR_HIP 3   /* a HIP runtime call failed                                                */
#define AC_ERR_STATE 4 /* call out of order, e.g. ac_run before both graphs are set                 */
#define AC_ERR_IO 5    /* file could not be opened, is truncated or malformed                      */
#define AC_ERR_NODEV 6 /* no HIP device / the requested device does not exist                      */

#define AC_DATA 0
#define AC_PATTERN 1

/* -------------------------------------------------------------------------------------------------------------
 * Host-side graph (the reference's Graph, src/main.cu:8-19, with FP32 features and int32 ids).
 * Arrays are row-major: node_feats[num_nodes * node_feat_dim], edge_feats[num_edges * edge_feat_dim].
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct ac_host_graph {
  int64_t num_nodes;
  int64_t node_feat_dim;
  int64_t num_edges;
  int64_t edge_feat_dim;
  float* node_feats;
  float* edge_feats;
  int32_t* srcs;
  int32_t* dsts;
} ac_host_graph;

/* Replaces loadGraph (src/main.cu:21-63): reads the binary layout written by prob2bin.py:46-64
 * (uint64 {nV, dV, nE, dE}, f64 vertex features, f64 edge features, uint64 srcs, uint64 dsts).
 * Unlike the reference it validates sizes and ids.  Arrays are malloc'd; release with ac_graph_free. */
int ac_graph_load_bin(const char* path, ac_host_graph* out);
/* Writes the same binary layout (the output of prob2bin.py:46-64). */
int ac_graph_save_bin(const char* path, const ac_host_graph* g);
/* Reads the CSV pair prob2bin.py:11-39 consumes: header line "N C" skipped; vertex rows "id f1..fd" (id must
 * equal the row index, prob2bin.py:28-29); edge rows "src dst f1..fd".  Whitespace separated, CRLF tolerated. */
int ac_graph_load_csv(const char* vertex_csv, const char* edge_csv, ac_host_graph* out);
/* Writes that CSV pair (values printed with 17 significant digits, so a float round-trips exactly). */
int ac_graph_save_csv(const char* vertex_csv, const char* edge_csv, const ac_host_graph* g);
void ac_graph_free(ac_host_graph* g);

/* The reference's Graph at its own precision (src/main.cu:8-19 with Int = uint64_t, Real = double,
 * src/helpers.hxx:3-4): what loadGraph reads and prob2bin.py writes, nothing rounded. */
typedef struct ac_graph64 {
  uint64_t num_nodes;
  uint64_t node_feat_dim;
  uint64_t num_edges;
  uint64_t edge_feat_dim;
  double* node_feats;
  double* edge_feats;
  uint64_t* srcs;
  uint64_t* dsts;
} ac_graph64;
/* Replaces loadGraph's file read (src/main.cu:21-47): the .bin as stored, validated (file length, ids < nV). */
int ac_graph64_load_bin(const char* path, ac_graph64* out);
/* Writes the .bin layout of prob2bin.py:46-64 (uint64 header, f64 features, uint64 srcs, uint64 dsts). */
int ac_graph64_save_bin(const char* path, const ac_graph64* g);
/* Reads the CSV pair as prob2bin.py:11-39 does, converting each decimal to the double pandas.read_csv produces by
 * default (which is not always the correctly rounded one), so ac_graph64_save_bin then writes the bytes
 * prob2bin.py writes. */
int ac_graph64_load_csv(const char* vertex_csv, const char* edge_csv, ac_graph64* out);
void ac_graph64_free(ac_graph64* g);

/* -------------------------------------------------------------------------------------------------------------
 * Synthetic inputs (no counterpart in the reference, whose datasets are absent; these make its configurations).
 * All generators are deterministic functions of their seed and use every host core.
 * ------------------------------------------------------------------------------------------------------------- */
/* Directed Erdős–Rényi G(n, m): m distinct (src, dst) pairs, no self-loops, sorted by (src, dst). */
int ac_gen_er(int64_t n, int64_t m, uint64_t seed, int32_t* srcs, int32_t* dsts);
/* R-MAT with 2^scale vertices and edge_factor * 2^scale draws, quadrant probabilities (a, b, c, 1-a-b-c);
 * self-loops and duplicates removed; sorted by (src, dst).  *srcs / *dsts are malloc'd (free with ac_free). */
int ac_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int64_t* num_edges,
                int32_t** srcs, int32_t** dsts);
/* out[i] = sigma * z_i, z_i ~ N(0,1) from a counter-based generator keyed by (seed, stream, i). */
int ac_gen_normal(float* out, int64_t count, uint64_t seed, uint64_t stream, float sigma);
/* Planted pattern: a connected subgraph of the data graph with n_pv vertices and n_pe of the edges they induce.
 * Outputs pattern edges (p_srcs/p_dsts in pattern ids), vmap[p] = data vertex of pattern vertex p and
 * emap[q] = data edge index of pattern edge q.  Returns AC_ERR_ARG if no such subgraph was found. */
int ac_gen_pattern(int64_t n, int64_t m, const int32_t* srcs, const int32_t* dsts, int n_pv, int n_pe,
                   uint64_t seed, int32_t* p_srcs, int32_t* p_dsts, int32_t* vmap, int64_t* emap);
/* k planted patterns at once (the data graph's incidence lists are built once; patterns planted in parallel):
 * pattern i has n_pv[i] vertices, n_pe[i] edges and seed seeds[i]; its outputs start at the sum of the preceding
 * counts in p_srcs / p_dsts / emap (edges) and vmap (vertices).  status[i] = 0 if planted, 1 if no subgraph with
 * that many induced edges was found. */
int ac_gen_patterns(int64_t n, int64_t m, const int32_t* srcs, const int32_t* dsts, int k, const int32_t* n_pv,
                    const int32_t* n_pe, const uint64_t* seeds, int32_t* p_srcs, int32_t* p_dsts, int32_t* vmap,
                    int64_t* emap, int32_t* status);
void ac_free(void* p);

/* -------------------------------------------------------------------------------------------------------------
 * Matcher context: the driver's buffers and its prep + loop (src/main.cu:80-247) on one GPU.
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct ac_ctx ac_ctx;

int ac_create(ac_ctx** out, int device);
int ac_destroy(ac_ctx* ctx);
const char* ac_last_error(const ac_ctx* ctx);
const char* ac_global_error(void);

/* Replaces the H2D half of loadGraph (src/main.cu:49-62) for which = AC_DATA or AC_PATTERN.  Host arrays are copied;
 * the context keeps the device copies plus the CSR/CSC views it builds from them.  dV/dE of the two graphs must
 * agree (cdist compares them, src/ac.hxx:33-48). */
int ac_set_graph(ac_ctx* ctx, int which, int64_t num_nodes, int64_t node_feat_dim, const float* node_feats,
                 int64_t num_edges, int64_t edge_feat_dim, const float* edge_feats, const int32_t* srcs,
                 const int32_t* dsts);
/* The same from the reference's own Graph (src/main.cu:49-62 uploads it unchanged): ids are checked and stored as
 * int32, features are rounded to FP32 for the device (the matcher's storage precision). */
int ac_set_graph64(ac_ctx* ctx, int which, const ac_graph64* g);

/* Runs prep (src/main.cu:107-148) and n_iter iterations of the loop (src/main.cu:160-245); n_iter < 0 means the
 * number of pattern vertices, as src/main.cu:160.  Synchronous. */
int ac_run(ac_ctx* ctx, int n_iter);
/* Enqueues the same work on the context's stream and returns without waiting (for timing loops). */
int ac_run_async(ac_ctx* ctx, int n_iter);
int ac_sync(ac_ctx* ctx);

/* Replaces ac::transpose + the D2H copy (src/main.cu:247,255-256): out[v * pn + p] = MU[p][v]. */
int ac_get_mu(ac_ctx* ctx, float* host_out);
/* Thresholded match matrix (no counterpart in the reference, which prints MU only): out[v * pn + p] = 1 if v is
 * the first data vertex attaining the column maximum of MU[.][p], or 2 if additionally MU[v][p] >= log_tau;
 * 0 otherwise. */
int ac_get_match(ac_ctx* ctx, float log_tau, uint8_t* host_out);
/* Top-k data vertices per pattern vertex (no counterpart in the reference, which prints MU only): for each p, the k
 * data vertices of largest MU[v][p], most likely first, ties by lower v: idx_out[p * k + r] (and, if val_out is not
 * NULL, val_out[p * k + r] = MU[v][p]).  1 <= k <= dn. */
int ac_get_topk(ac_ctx* ctx, int k, int32_t* idx_out, float* val_out);
/* Sizes {dn, de, pn, pe}. */
int ac_get_sizes(const ac_ctx* ctx, int64_t* out4);
/* Device time of the last run, from HIP events: prep (src/main.cu:107-148) and loop (src/main.cu:160-245). */
int ac_get_timing(const ac_ctx* ctx, double* prep_ms, double* loop_ms);
/* No counterpart in the reference: {edge bundles, bundles whose costs the last run's walk read as 24-bit codes}
 * (AC_OPT_PACK_COSTS); synchronises the context's stream. */
int ac_get_cost_packing(ac_ctx* ctx, int64_t* out2);
/* Device stream the context runs on (a hipStream_t).  The prep forks part of its work onto a second stream of the
 * context and joins it back with events, so work ordered after a run on this stream sees all of the run. */
void* ac_stream(ac_ctx* ctx);

/* Stages of one run, timed separately by ac_run_profiled (they replace the reference's NVTX ranges "prep", "loop",
 * "step1".."step5", src/main.cu:107,161-244). */
#define AC_STAGE_PREP 0        /* cdist, edge-cost statistics, prep maxima (src/main.cu:107-148)              */
#define AC_STAGE_VERTEX_A 1    /* VF / VR, column maxima, edge log-sum-exps (steps 1, 3a, 4a)                 */
#define AC_STAGE_REDUCE_A 2    /* combine the per-block partials of stage 1                                   */
#define AC_STAGE_EDGE_WALK 3   /* destination maxima over the data edges (steps 2, 4b, 4c)                    */
#define AC_STAGE_WALK_FIXUP 4  /* combine the pieces of vertices with very long in-edge lists                 */
#define AC_STAGE_VERTEX_C 5    /* MU column log-sum-exp (step 5)                                               */
#define AC_STAGE_REDUCE_C 6    /* combine the per-block partials of stage 5                                   */
#define AC_STAGE_OUTPUT 7      /* normalised MU written in dn x pn order (src/main.cu:247)                     */
#define AC_NSTAGES 8
/* Like ac_run, with a HIP event after every launch; stage_ms[s] (AC_NSTAGES entries) receives the device time
 * of stage s summed over the iterations.  Extra events make it slightly slower than ac_run. */
int ac_run_profiled(ac_ctx* ctx, int n_iter, double* stage_ms);

/* -------------------------------------------------------------------------------------------------------------
 * Tuning options.  No counterpart in the reference (its launch shapes are fixed); used by the autotuner below and
 * by callers that pin a configuration.
 * ------------------------------------------------------------------------------------------------------------- */
#define AC_OPT_WALK_BATCH 1  /* edge-walk steps per software-pipelined batch: 1, 2 or 4 (0 = default, 4)         */
#define AC_OPT_GRAPH 2       /* 1: replay iteration pairs from a captured HIP graph (fewer launches); 0: launch each */
#define AC_OPT_SLAB_EDGES 3  /* batched runs: most pattern edges packed into one slab, 4..256 (default 128)         */
#define AC_OPT_TAPE 4        /* > 0: runs record what ac_backward needs, checkpointing every this many iterations;
                              * 1: every state, and every walk's argmax rows too, so ac_backward recomputes nothing */
#define AC_OPT_PRUNE 5       /* 1 (default): the edge walk skips bundles of edges whose costs are all at or above a
                              * per-pattern-edge bound past which no edge term can exceed the fill value, so the
                              * result is bit-identical; 0: every edge is walked                                  */
#define AC_OPT_CDIST_MFMA 6  /* 1 (default): feature distances on FP64 matrix cores in GEMM form, pairs that lose
                              * digits to cancellation recomputed directly; 0: direct form on the vector units     */
#define AC_OPT_PACK_COSTS 7  /* 1 (default): the iteration's edge walk reads the edge costs of a bundle as 24-bit codes
                              * (bit pattern minus a per-bundle exponent base) when all of them lie within two
                              * binades, losslessly, so results are bit-identical; 0: FP32 costs throughout        */
#define AC_OPT_CNULL 8        /* 1: the PNNL Cnull term, which the reference leaves out (src/main.cu:83, test.py:97):
                              * every iteration's fills become VRmax - cnull (FMax) and VFmax - cnull (RMax), with
                              * cnull the log-softmax over the pattern's edges of max(||edge features||_2,
                              * max_e normprob(CE)) (test.py:93-95,148-149); 0 (default): left out, as the reference.
                              * Not differentiated: runs with AC_OPT_TAPE > 0 reject it                            */
int ac_set_option(ac_ctx* ctx, int option, int value);
int ac_get_option(const ac_ctx* ctx, int option, int* value);

/* -------------------------------------------------------------------------------------------------------------
 * Adjoint (no counterpart in the reference): per-dimension cdist weights, d = sqrt(sum_k w_k (a_k - b_k)^2), and
 * the gradient of a loss on the output MU with respect to them, by reverse mode through the prep and every
 * iteration (max subgradients routed through the argmax rows / vertices recorded in the forward pass; the
 * log-softmax Jacobians in FP32), recomputing each segment of AC_OPT_TAPE iterations from its checkpoint (with
 * AC_OPT_TAPE = 1 the run has recorded every state and nothing is recomputed).
 * ------------------------------------------------------------------------------------------------------------- */
/* Weights of the vertex (dV) and edge (dE) feature distances, host arrays, >= 0; NULL = unweighted (the
 * reference's cdist, src/ac.hxx:33-56).  Set after the data graph. */
int ac_set_weights(ac_ctx* ctx, const float* wv, const float* we);
/* After ac_run with AC_OPT_TAPE > 0: grad_mu = dL/dMU (dn x pn, the ac_get_mu layout, host) -> grad_wv = dL/dwv
 * (dV), grad_we = dL/dwe (dE), host arrays (either may be NULL). */
int ac_backward(ac_ctx* ctx, const float* grad_mu, float* grad_wv, float* grad_we);

/* -------------------------------------------------------------------------------------------------------------
 * Batched scoring: many patterns against the context's resident data graph.  No counterpart in the reference,
 * whose driver scores one pattern per process (src/main.cu:66-257); each pattern's result equals that of a plain
 * run of the same pattern (ac_set_graph + ac_run) to within rounding of the log-sum-exp reductions.
 *
 * Patterns are packed into slabs (disjoint unions of patterns: the matcher's normalisations are per pattern vertex
 * and per pattern edge, so patterns in one slab do not interact), longest iteration count first; each slab runs the
 * fused pipeline once, and each pattern's MU is copied out after its own iteration count.  A batched run replaces
 * the context's current pattern.
 * ------------------------------------------------------------------------------------------------------------- */
/* k patterns, concatenated: pattern i has vertices [v_off[i], v_off[i+1]) of vfeat (dV features each) and edges
 * [e_off[i], e_off[i+1]) of efeat / src / dst, whose endpoints are local vertex ids in [0, v_off[i+1] - v_off[i]).
 * Each pattern has 1..256 vertices and at most 256 edges.  Host arrays, copied. */
int ac_batch_set_patterns(ac_ctx* ctx, int k, const int32_t* v_off, const int32_t* e_off, int64_t dV,
                          const float* vfeat, int64_t dE, const float* efeat, const int32_t* src, const int32_t* dst);
/* Runs every pattern: n_iter[i] iterations for pattern i (NULL: its vertex count, as src/main.cu:160). */
int ac_batch_run(ac_ctx* ctx, const int32_t* n_iter);
/* MU of pattern i after ac_batch_run: dn x pn_i, data-vertex-major (src/main.cu:247 order). */
int ac_batch_get_mu(ac_ctx* ctx, int i, float* host_out);
/* out[0] = patterns, out[1] = slabs, out[2] = sum over patterns of de * pe_i * iterations_i (edge-pair updates),
 * out[3] = sum over slabs of de * slab edges * slab iterations (updates computed, including padding). */
int ac_batch_info(const ac_ctx* ctx, int64_t* out4);
/* Device time of the last ac_batch_run: prep of all slabs, and iterations + outputs of all slabs. */
int ac_batch_timing(const ac_ctx* ctx, double* prep_ms, double* loop_ms);
/* Picks AC_OPT_SLAB_EDGES and AC_OPT_WALK_BATCH for the patterns set by ac_batch_set_patterns: for each candidate
 * slab size, the slab with the most iterations is prepared once and `probe_iters` iterations of it are timed with
 * each walk batch; the candidate with the least estimated total time of all slabs is set.  choice_out (may be NULL)
 * receives {slab edges, walk batch}; est_ms_out (may be NULL) the estimated time. */
int ac_batch_autotune(ac_ctx* ctx, int probe_iters, int32_t* choice_out, double* est_ms_out);

/* -------------------------------------------------------------------------------------------------------------
 * Stage operators on caller-owned device buffers: the ac:: operator library (src/ac.hxx) with FP32 values and
 * int32 ids, same argument meaning and pattern-major layouts.  Asynchronous on `stream`.
 * ------------------------------------------------------------------------------------------------------------- */
/* ac::cdist (src/ac.hxx:33-56): out[j * n_b + i] = || a_j - b_i ||_2 for j < n_a, i < n_b. */
int ac_cdist(int64_t n_a, int64_t n_b, int64_t dim, const float* feats_a, const float* feats_b, float* out,
             void* stream);
/* ac::RowSoftmax2 (src/ac.hxx:77-113): in-place log-softmax of each row of an n_row x n_col matrix, computed
 * max-stabilised: x - max - log1p(sum_{k != argmax} exp(x_k - max)). */
int ac_row_log_softmax(int64_t n_row, int64_t n_col, float* x, void* stream);
/* No counterpart in the reference: ac_row_log_softmax for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The same kernels and results (bit for bit), with
 * the partials in a fixed device-resident scratch of 1 MiB per GPU instead of a stream-ordered allocation; it calls
 * no allocating routine.  At most 32768 rows (AC_ERR_ARG beyond).  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams or threads) race and must be ordered by the caller. */
int ac_row_log_softmax_static(int64_t n_row, int64_t n_col, float* x, void* stream);
/* ac::RowMax2 (src/ac.hxx:58-75): out[r] = max_c in[r * n_col + c]. */
int ac_row_max(int64_t n_row, int64_t n_col, const float* in, float* out, void* stream);
/* ac::EdgeMaxReduce2_t (src/ac.hxx:115-149): xmax[r * n_col_out + v] = max(vymax[r],
 * max_{c : nodes[c] = v} xe[r * n_col_in + c]).  Exact (bit-identical for any execution order). */
int ac_edge_max_reduce(int64_t n_col_in, int64_t n_col_out, int64_t n_row, const float* vymax, const float* xe,
                       float* xmax, const int32_t* nodes, void* stream);
/* ac::ComputeMU2_t (src/ac.hxx:151-186): mu[p * n_col_out + v] = -cv[p * n_col_out + v]
 *   + sum_{q : dsts[q] = p} fmax[q * n_col_in + v] + sum_{q : srcs[q] = p} rmax[q * n_col_in + v],
 * summed in ascending q (fmax term before rmax term), so the result is deterministic. */
int ac_compute_mu(int64_t n_col_in, int64_t n_row_in, int64_t n_col_out, int64_t n_row_out, const float* cv,
                  const float* fmax, const float* rmax, const int32_t* srcs, const int32_t* dsts, float* mu,
                  void* stream);
/* ac::transpose (src/ac.hxx:17-31): out[c * n_rows + r] = in[r * n_cols + c]. */
int ac_transpose(const float* in, float* out, int64_t n_rows, int64_t n_cols, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AC_MI355X_H */
