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
  int fix_shift;       // headroom of the fixed-point source sums: the largest out-degree is at most 2^(23 + fix_shift)
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
// The FMax edge terms of one step back: the row FA[v][q] recorded for each (v, q) whose edge term carries gradient
// gE = gEF[v][q] != 0 receives -gE in gCE, and its source's sum gVRfix[src][q] gains gE in fixed point (scale from the
// step's largest |gE_F|, gmax).
struct FmaxTerms {
  const float* gEF;             // [n][Qs] the step's gE_F
  const int* FA;                // [n][Qs] the state's FMax argmax rows
  const int* bsrc;              // [brows] the source of each bundle row (bit 31: the run's vertex flag)
  int64_t brows;
  float* gCE;                   // [brows][Qs]
  unsigned long long* gVRfix;   // [n][Qs] fixed-point source sums (two's complement), zero between steps
  const unsigned* gmax;         // the step's largest |gE_F| as float bits (k_bwd_route)
};

// gVF / gVR: those of the next state (the previous step back; zeros at the first), whose negations are this state's
// FMax / RMax gradients.  gmax (nullable, zero before): receives the largest |gE_F|; gERsum (nullable) accumulates
// gE_R over the steps.
void launch_bwd_route(const BwdGraph& G, const BwdState& S, const float* gZ, int with_gz, const float* gVF,
                      const float* gVR, float* gEF, float* gER, unsigned* gmax, float* gERsum, double* part,
                      int nparts, double* red4, hipStream_t st);
// Also the step's FMax edge terms FE.
void launch_bwd_in(const BwdGraph& G, const BwdState& S, const double* MU, const QStat* qs1, const double* red4,
                   const float* gER, float* gVR, float* gVF, double* cD, const FmaxTerms& FE, hipStream_t st);
// The prep's FMax edge terms: gCE only (its state has no VR).
void launch_bwd_fscatter(const BwdGraph& G, const float* gEF, const int* FA, const int* bsrc, int64_t brows,
                         float* gCE, hipStream_t st);
// RMax edge terms, once after the last step: gCE[AO[v][q]][q] -= gERsum[v][q].
void launch_bwd_rterms(const BwdGraph& G, const float* gERsum, const int* AO, int64_t brows, float* gCE,
                       hipStream_t st);
// The vertex side of a step back (k_bwd_vtx): adds the fixed-point sums gVRfix (scale from gmax) to gVR and clears
// them, writes gMU, and leaves bwd_vtx_blocks(G) partial column sums of gMU in part ([blocks][P]).
int bwd_vtx_blocks(const BwdGraph& G);
void launch_bwd_vtx(const BwdGraph& G, float* gVR, const float* gVF, unsigned long long* gVRfix, const unsigned* gmax,
                    float* gMU, double* part, hipStream_t st);
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
