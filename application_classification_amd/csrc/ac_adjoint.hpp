// Adjoint kernels (ac_adjoint.hip): device views and launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ac_internal.hpp"

namespace ac {

// The fixed inputs of the vertex side: costs, out-edge statistics and the pattern's incidence lists.
struct BwdGraph {
  int64_t n;
  int P, Q, Qs;
  const float* CV;      // [n][P]
  const float* minCEo;  // [n][Qs]
  const float* lsO;     // [n][Qs]
  const int* ps;
  const int* pd;
  const int* inc_off;  // [P + 1]
  const int* inc;      // (q << 1) | (1 if q starts at p), ascending q
};

// One state of the iteration: FMaxN / VFm rows, its scalars and the log-sum-exp of its MU columns.
struct BwdState {
  const float* FMaxN;  // [n][Qs]
  const float* VFm;    // [n][Qs]
  const QStat* qs;     // [Q]
  const Lse2* pst;     // [P]
  int first;           // the prep state: MU = LSM(-CV), no FMax / RMax terms
};

void launch_fill_int(int* p, int64_t count, int value, hipStream_t st);
void launch_sum_parts(int nparts, int C, const double* part, double* out, hipStream_t st);
void launch_colsum(int64_t n, int C, const float* X, double* part, int nparts, double* out, hipStream_t st);
void launch_bwd_out(const BwdGraph& G, const double* MU, const double* colsum, float* gMU, float* gCV, hipStream_t st);
// The argmax masks of the m CSR positions: bit q of word [k][q / 32] marks that the row at position k is the FMax
// argmax of column q of its destination and that edge term carries gradient.  k_bwd_route sets them from FA and the
// gather reads and clears them, so they are zero between uses (and must be zero at the first).
struct ArgMasks {
  const int* FA;      // [n][Qs] the state's FMax argmax rows
  const int* rowpos;  // [brows] the CSR position of each bundle row (-1: padding)
  int64_t brows, m;
  int W;              // words per position: ceil(Qs / 32)
  unsigned* bits;     // [m][W] (nullptr: no masks)
};

// gVF / gVR: those of the next state (the previous step back; zeros at the first), whose negations are this state's
// FMax / RMax gradients.  Sets the argmax masks am; gERsum (nullable) accumulates gE_R over the steps.
void launch_bwd_route(const BwdGraph& G, const BwdState& S, const float* gZ, int with_gz, const float* gVF,
                      const float* gVR, float* gEF, float* gER, const ArgMasks& am, float* gERsum, double* part,
                      int nparts, double* red4, hipStream_t st);
void launch_bwd_in(const BwdGraph& G, const BwdState& S, const double* MU, const QStat* qs1, const double* red4,
                   const float* gER, float* gVR, float* gVF, double* cD, hipStream_t st);
// FMax edge terms of one iteration (or of the prep, gVR == nullptr), gathered over the CSR runs `runs` (pairs ks =
// {bundle row, source}, kdst = the destination at each CSR position) from the argmax masks that k_bwd_route set, which
// it clears; carry: one [Qs] FP64 row per carry slot of the runs.  Deterministic.
void launch_bwd_gather(const BwdGraph& G, int64_t nruns, const Run* runs, int nsplit, const Split* splits,
                       const int2* ks, const int* kdst, const float* gEF, float* gVR, float* gCE, double* carry,
                       unsigned* argbits, hipStream_t st);
// RMax edge terms, once after the last step: gCE[AO[v][q]][q] -= gERsum[v][q].
void launch_bwd_rterms(const BwdGraph& G, const float* gERsum, const int* AO, int64_t brows, float* gCE,
                       hipStream_t st);
// out[k] = bdst[ks[k].x]: the destination of the edge at each of the m CSR positions; rowpos[row] = the CSR position
// of each bundle row (-1: padding).
void launch_csr_dst(int64_t m, const int2* ks, const int* bdst, int* out, int* rowpos, int64_t brows,
                    hipStream_t st);
void launch_bwd_mu(const BwdGraph& G, const float* gVR, const float* gVF, float* gMU, hipStream_t st);
void launch_bwd_prep_fill(const BwdGraph& G, const double* red4, const int* am0, float* gMU, hipStream_t st);
void launch_argmin_out(int64_t m, int Q, int Qs, const int2* ks, const float* CE, const float* minCEo, int* AO,
                       hipStream_t st);
void launch_col_argmin(int64_t n, int P, const float* CV, float* pv, int* pi, int nparts, int* am, hipStream_t st);
void launch_bwd_edge_coef(int64_t nrows, int Q, int Qs, const int* beid, const int* bsrc, const float* CE,
                          const double* cD, const double* cq, float* G, hipStream_t st);
void launch_bwd_vertex_coef(int64_t total, const float* CV, float* G, hipStream_t st);
void launch_outer_acc(int64_t nrows, const int* idx, const float* feats, int dim, const float* G, int Gs, int K,
                      double* part, int nparts, hipStream_t st);

}  // namespace ac
