// Adjoint of the matcher (no counterpart in the reference): gradients of a loss on the output log-probabilities with
// respect to the per-dimension cdist weights, by reverse mode through the prep and every iteration.
//
// One iteration maps the state S_i = (MU_i, FMax_i, RMax_i) to S_{i+1} (src/main.cu:160-245):
//   VF = MU_i[pd] - FMax_i,  VR = MU_i[ps] - RMax_i,  fillR = max_v VF,  fillF = max_v VR
//   FMax_{i+1}[v] = max(fillF, max_{e->v} (VR[src e] - CE_e) - LSE_FE),  LSE_FE = log sum_e exp(VR[src e] - CE_e)
//   RMax_{i+1}[v] = max(fillR, max_{e: src=v} (VF[v] - CE_e) - LSE_RE)
//   MU_{i+1} = LSM_v(-CV + sum_{q: pd=p} FMax_{i+1} + sum_{q: ps=p} RMax_{i+1})
// Its adjoint, given (gMU, gFMax, gRMax) of S_{i+1}:
//   gZ = gMU - exp(MU_{i+1}) colsum_v(gMU)                        (log-softmax Jacobian, FP32)
//   gCV -= gZ;  gF = gFMax + gZ[pd],  gR = gRMax + gZ[ps]
//   max subgradients: to the recorded argmax in-edge (FMax) / the minimum-cost out-edge (RMax) when the edge term
//   wins, else to the fill, whose own subgradient goes to the recorded argmax vertex;
//   log-sum-exp over edges: sum_e exp(FE_e) routes -S_F = -sum_v gE_F to every edge; its vertex part is the closed
//   form  sum_{e: src=v} exp(VR[v] - CE_e - LSE_FE) = exp(VR[v] - LSE_FE - minCEo[v] + lsO[v]), and its edge-cost
//   part S_F exp(VR[src] - LSE_FE) exp(-CE_e) is accumulated per (src, q) in cD, so that the final gradient of every
//   edge cost is  gCE_e = (sparse argmax terms) + (cD[src e] + c_q) exp(-CE_e), formed once at the end.
// The weight gradients are then  dL/dw_k = sum gC / (2 C) (x_k - p_k)^2  over vertex and edge costs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "ac_adjoint.hpp"

namespace ac {

namespace {

constexpr double kInf = __builtin_huge_val();
typedef float f4v __attribute__((ext_vector_type(4)));

// MU[v][p] of a state: (Z - m) - l with Z = -CV + (first ? 0 : the incidence terms of p in ascending q, the FMax
// term before the RMax term), exactly as the vertex pass evaluates it.
__device__ __forceinline__ double mu_of(const BwdGraph& G, const BwdState& S, int64_t v, int p) {
  double acc = -(double)G.CV[v * G.P + p];
  if (!S.first)
    for (int k = G.inc_off[p]; k < G.inc_off[p + 1]; ++k) {
      const int code = G.inc[k];
      const int q = code >> 1;
      const QStat s = S.qs[q];
      if (code & 1) {
        const float vfm = S.VFm[v * G.Qs + q];
        const double rn = vfm == -INFINITY ? -kInf : ((double)vfm - s.mRE) - s.lRE;
        acc += fmax(s.fillR, rn);
      } else {
        acc += fmax(s.fillF, (double)S.FMaxN[v * G.Qs + q]);
      }
    }
  return (acc - S.pst[p].m) - S.pst[p].l;
}

__device__ __forceinline__ double fmax_of(const BwdGraph& G, const BwdState& S, int64_t v, int q) {
  return fmax(S.qs[q].fillF, (double)S.FMaxN[v * G.Qs + q]);
}

__device__ __forceinline__ double rn_of(const BwdGraph& G, const BwdState& S, int64_t v, int q) {
  const float vfm = S.VFm[v * G.Qs + q];
  return vfm == -INFINITY ? -kInf : ((double)vfm - S.qs[q].mRE) - S.qs[q].lRE;
}

}  // namespace

// Column sums of X[n][C] (C <= 256): per-block partials in a fixed order.
__global__ __launch_bounds__(256) void k_colsum(int64_t n, int C, const float* __restrict__ X, double* __restrict__ part) {
  __shared__ double red[256];
  const int t = threadIdx.x, RC = 256 / C;
  const bool active = t < RC * C;
  const int c = active ? t % C : 0;
  double s = 0.0;
  if (active)
    for (int64_t v = (int64_t)blockIdx.x * RC + t / C; v < n; v += (int64_t)gridDim.x * RC) s += (double)X[v * C + c];
  red[t] = s;
  __syncthreads();
  if (t < C) {
    double a = 0.0;
    for (int r = 0; r < RC; ++r) a += red[r * C + t];
    part[(int64_t)blockIdx.x * C + t] = a;
  }
}

// out[c] = sum over parts of part[b][c]: a block takes CB = min(C, 32) columns and 256 / CB lanes per column, lane l
// summing parts l, l + L, ... in order, the lanes then summed in order (deterministic).
__global__ __launch_bounds__(256) void k_sum_parts(int nparts, int C, const double* __restrict__ part,
                                                   double* __restrict__ out) {
  // One block per column: lane l sums partials l, l + 256, ... in order (eight loads in flight before the adds), then
  // a fixed tree over the lanes.  The order is the same on every call.
  __shared__ double red[256];
  const int t = threadIdx.x, c = blockIdx.x;
  double a = 0.0;
  for (int b0 = t; b0 < nparts; b0 += 8 * 256) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int b = b0 + u * 256;
      v[u] = part[(int64_t)(b < nparts ? b : nparts - 1) * C + c];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (b0 + u * 256 < nparts) a += v[u];
  }
  red[t] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) out[c] = red[0];
}

// Log-softmax Jacobian of MU = LSM_v(Z): gMU <- gZ = gMU - exp(MU) colsum(gMU); gCV -= gZ.
__global__ __launch_bounds__(256) void k_bwd_out(int64_t total, int P, const double* __restrict__ MU,
                                                 const double* __restrict__ colsum, float* __restrict__ gMU,
                                                 float* __restrict__ gCV) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const float gz = gMU[i] - __expf((float)MU[i]) * (float)colsum[i % P];
    gMU[i] = gz;
    gCV[i] -= gz;
  }
}

// Max-subgradient routing of FMax / RMax of state S (thread per (v, q), q fixed per thread).  gF = -gVF + gZ[pd]
// (with_gz), gR = -gVR + gZ[ps], where gVF / gVR are the next state's (the step processed before this one; zero at the
// first): the gradients of this state's FMax / RMax through the next step's inputs.  The edge term takes the
// gradient where it is strictly above the fill.  Per-block partials part[b][4][Q]: sum gE_F, sum gE_R, fill gradient
// of fillF, of fillR.  gmax (nullable) receives the largest |gE_F| (as float bits, an atomic maximum: exact and
// order-free), which sets the fixed-point scale of the FMax edge terms' source sums.  gERsum (nullable) accumulates
// gE_R per (v, q) over the steps: the RMax rows AO are the same in every step, so their edge terms are applied once.
__global__ __launch_bounds__(256) void k_bwd_route(BwdGraph G, BwdState S, const float* __restrict__ gZ, int with_gz,
                                                   const float* __restrict__ gVF, const float* __restrict__ gVR,
                                                   float* __restrict__ gEF, float* __restrict__ gER,
                                                   unsigned* __restrict__ gmax, float* __restrict__ gERsum,
                                                   double* __restrict__ part) {
  __shared__ double red[256][4];
  __shared__ unsigned bmax;
  const int t = threadIdx.x, Q = G.Q, RQ = 256 / Q;
  if (t == 0) bmax = 0u;
  __syncthreads();
  const bool active = t < RQ * Q;
  const int q = active ? t % Q : 0;
  double sF = 0.0, sR = 0.0, fF = 0.0, fR = 0.0;
  unsigned mx = 0u;
  if (active) {
    const int pdq = G.pd[q], psq = G.ps[q];
    const double fillF = S.qs[q].fillF, fillR = S.qs[q].fillR;
    for (int64_t v = (int64_t)blockIdx.x * RQ + t / Q; v < G.n; v += (int64_t)gridDim.x * RQ) {
      const int64_t i = v * G.Qs + q;
      float gF = -gVF[i], gR = -gVR[i];
      if (with_gz) {
        gF += gZ[v * G.P + pdq];
        gR += gZ[v * G.P + psq];
      }
      const bool eF = (double)S.FMaxN[i] > fillF;
      const bool eR = rn_of(G, S, v, q) > fillR;
      gEF[i] = eF ? gF : 0.f;
      gER[i] = eR ? gR : 0.f;
      if (eF) mx = max(mx, __float_as_uint(fabsf(gF)));
      if (gERsum && eR && gR != 0.f) gERsum[i] += gR;
      if (eF) sF += gF; else fF += gF;
      if (eR) sR += gR; else fR += gR;
    }
  }
  if (gmax && mx) atomicMax(&bmax, mx);
  red[t][0] = sF;
  red[t][1] = sR;
  red[t][2] = fF;
  red[t][3] = fR;
  __syncthreads();
  if (gmax && t == 0 && bmax) atomicMax(gmax, bmax);
  if (t < Q)
    for (int k = 0; k < 4; ++k) {
      double a = 0.0;
      for (int r = 0; r < RQ; ++r) a += red[r * Q + t][k];
      part[((int64_t)blockIdx.x * 4 + k) * Q + t] = a;
    }
}

// The fixed-point scale of the FMax edge terms' source sums: 2^(39 - shift - e) for the largest |gE_F| in
// [2^e, 2^(e+1)), so every term is below 2^(40 - shift) in magnitude.  A source's sum has at most its out-degree of
// terms, and shift (BwdGraph::fix_shift) leaves room for the graph's largest out-degree: up to 2^(23 + shift) terms
// stay inside int64.  A power of two, so scaling is exact and only the rounding to an integer (at 2^-(40 - shift)
// of the largest term) loses anything.
__device__ __forceinline__ int fix_exponent(unsigned gmax_bits, int shift) {
  const int e = (int)((gmax_bits >> 23) & 0xFFu) - 127;  // subnormal maxima: e = -127, a larger scale than needed
  return 39 - shift - max(-126, min(e, 127));
}

// The FMax edge term of (v, q), where it carries gradient gE = gEF[v][q] != 0: its recorded row FA[v][q] receives
// -gE in gCE (one term per (row, q) and step, so the float atomic adds the same two values in the same order as a
// plain update), and the row's source accumulates gE in the fixed-point sum gVRfix[src][q] (integer additions: exact,
// so the sum does not depend on their order).  FA rows exist only where the walk recorded a maximum (vertices with
// in-edges, bundles not pruned); a nonzero gradient implies one, and the row is bounds-checked besides.
__device__ __forceinline__ void fmax_edge_term(int64_t o, int q, int Qs, float ge, const int* __restrict__ FA,
                                               const int* __restrict__ bsrc, int64_t brows, float* __restrict__ gCE,
                                               unsigned long long* __restrict__ gVRfix, double fscale) {
  const int row = FA[o];
  if (row < 0 || row >= brows) return;
  unsafeAtomicAdd(gCE + (int64_t)row * Qs + q, -ge);
  if (gVRfix) {
    const int64_t src = bsrc[row] & ~kNewVertex;
    atomicAdd(gVRfix + src * Qs + q, (unsigned long long)__double2ll_rn((double)ge * fscale));
  }
}

// The prep's FMax edge terms (no source sums: the prep state has no VR).
__global__ void k_bwd_fscatter(int64_t n, int Q, int Qs, const float* __restrict__ gEF, const int* __restrict__ FA,
                               const int* __restrict__ bsrc, int64_t brows, float* __restrict__ gCE) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * Qs; i += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(i % Qs);
    if (q >= Q) continue;
    const float ge = gEF[i];
    if (ge != 0.f) fmax_edge_term(i, q, Qs, ge, FA, bsrc, brows, gCE, nullptr, 0.0);
  }
}

// Input side of one iteration (state S = S_i, scalars qs1 of S_{i+1}; red4 = {S_F, S_R, gfillF, gfillR}[Q]):
//   gVR = -S_F exp(VR - LSE_FE - minCEo + lsO) + [v = argmax VR] gfillF
//   gVF = gE_R - S_R exp(VF - LSE_RE - minCEo + lsO) + [v = argmax VF] gfillR
//   cD += S_F exp(VR - LSE_FE) + S_R exp(VF - LSE_RE)            (vertices with out-edges; FP64)
// and the FMax edge terms of the step (fmax_edge_term; FE: the step's gE_F, argmax rows and fixed-point scale), whose
// source sums go to gVRfix and are added to gVR by the vertex kernel that follows.
__global__ __launch_bounds__(256) void k_bwd_in(BwdGraph G, BwdState S, const double* __restrict__ MU,
                                                const QStat* __restrict__ qs1, const double* __restrict__ red4,
                                                const float* __restrict__ gER, float* __restrict__ gVR,
                                                float* __restrict__ gVF, double* __restrict__ cD, FmaxTerms FE) {
  const int Q = G.Q;
  const int64_t total = G.n * Q;
  const double fscale = ldexp(1.0, fix_exponent(*FE.gmax, G.fix_shift));
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = i / Q;
    const int q = (int)(i - v * Q);
    const int64_t o = v * G.Qs + q;
    const double VF = MU[v * G.P + G.pd[q]] - fmax_of(G, S, v, q);
    const double VR = MU[v * G.P + G.ps[q]] - fmax(S.qs[q].fillR, rn_of(G, S, v, q));
    const QStat s1 = qs1[q];
    const double SF = red4[q], SR = red4[Q + q];
    const float mo = G.minCEo[o];
    double gvr = 0.0, gvf = (double)gER[o];
    if (mo != INFINITY) {
      const double eF = exp(VR - (s1.mFE + s1.lFE)), eR = exp(VF - (s1.mRE + s1.lRE));
      const double wo = exp((double)G.lsO[o] - (double)mo);
      gvr -= SF * eF * wo;
      gvf -= SR * eR * wo;
      cD[o] += SF * eF + SR * eR;
    }
    if (v == s1.vF) gvr += red4[2 * Q + q];
    if (v == s1.vR) gvf += red4[3 * Q + q];
    gVR[o] = (float)gvr;
    gVF[o] = (float)gvf;
    const float ge = FE.gEF[o];
    if (ge != 0.f) fmax_edge_term(o, q, G.Qs, ge, FE.FA, FE.bsrc, FE.brows, FE.gCE, FE.gVRfix, fscale);
  }
}

// The vertex side of a step back, one tile of T vertices per block iteration:
//   1. gVR[v][q] += gVRfix[v][q] 2^-k, the fixed-point sums of the step's FMax edge terms (where nonzero; the
//      accumulator is cleared for the next step), staged with gVF in LDS;
//   2. gMU[v][p] = sum over the incidences of p (ascending q) of gVF (q ends at p) or gVR (q starts at p);
//   3. per-block partial column sums of gMU in FP64, vertices in order: part[b][p].
// The input state's FMax / RMax gradients are -gVF / -gVR, which k_bwd_route of the next step back reads directly.
__global__ __launch_bounds__(256) void k_bwd_vtx(BwdGraph G, float* __restrict__ gVR, const float* __restrict__ gVF,
                                                 unsigned long long* __restrict__ gVRfix,
                                                 const unsigned* __restrict__ gmax, float* __restrict__ gMU,
                                                 double* __restrict__ part, int T) {
  extern __shared__ float vsh[];  // [T][Qs] gVR, [T][Qs] gVF, [T][P] gMU, then the incidence lists
  const int Qs = G.Qs, P = G.P, t = threadIdx.x;
  float* tR = vsh;
  float* tF = tR + T * Qs;
  float* tM = tF + T * Qs;
  int* ioff = reinterpret_cast<int*>(tM + T * P);
  int* iinc = ioff + (P + 1);
  const int ninc = G.inc_off[P];
  for (int k = t; k <= P; k += 256) ioff[k] = G.inc_off[k];
  for (int k = t; k < ninc; k += 256) iinc[k] = G.inc[k];
  const double inv = ldexp(1.0, -fix_exponent(*gmax, G.fix_shift));
  double col = 0.0;  // thread t < P: this block's partial sum of column t
  __syncthreads();
  for (int64_t v0 = (int64_t)blockIdx.x * T; v0 < G.n; v0 += (int64_t)gridDim.x * T) {
    const int nv = (int)min((int64_t)T, G.n - v0);
    for (int e = t; e < nv * Qs; e += 256) {
      const int64_t o = v0 * Qs + e;
      float r = gVR[o];
      const long long f = (long long)gVRfix[o];
      if (f != 0) {
        r = (float)((double)r + (double)f * inv);
        gVR[o] = r;
        gVRfix[o] = 0ull;
      }
      tR[e] = r;
      tF[e] = gVF[o];
    }
    __syncthreads();
    for (int e = t; e < nv * P; e += 256) {
      const int vl = e / P, p = e - vl * P;
      float a = 0.f;
      for (int k = ioff[p]; k < ioff[p + 1]; ++k) {
        const int code = iinc[k];
        a += (code & 1) ? tR[vl * Qs + (code >> 1)] : tF[vl * Qs + (code >> 1)];
      }
      tM[e] = a;
      gMU[v0 * P + e] = a;
    }
    __syncthreads();
    if (t < P)
      for (int vl = 0; vl < nv; ++vl) col += (double)tM[vl * P + t];
  }
  if (t < P) part[(int64_t)blockIdx.x * P + t] = col;
}

// Fill subgradients of the prep state go to the argmax of MU0 = LSM(-CV) of the source / destination pattern vertex:
// fillF_q = max_v MU0[v][ps q], fillR_q = max_v MU0[v][pd q].  One thread, ascending q.
__global__ void k_bwd_prep_fill(BwdGraph G, const double* __restrict__ red4, const int* __restrict__ am0,
                                float* __restrict__ gMU) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = 0; q < G.Q; ++q) {
    const int ps = G.ps[q], pd = G.pd[q];
    gMU[(int64_t)am0[ps] * G.P + ps] += (float)red4[2 * G.Q + q];
    gMU[(int64_t)am0[pd] * G.P + pd] += (float)red4[3 * G.Q + q];
  }
}

// Rows of the minimum cost among out-edges: AO[src][q] = lowest CE row whose cost equals minCEo[src][q] (CSR order
// pairs {row, src}); AO starts at INT_MAX.
__global__ void k_argmin_out(int64_t m, int Q, int Qs, const int2* __restrict__ ks, const float* __restrict__ CE,
                             const float* __restrict__ minCEo, int* __restrict__ AO) {
  const int64_t total = m * Q;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i / Q;
    const int q = (int)(i - k * Q);
    const int2 e = ks[k];
    if (CE[(int64_t)e.x * Qs + q] == minCEo[(int64_t)e.y * Qs + q]) atomicMin(&AO[(int64_t)e.y * Qs + q], e.x);
  }
}

// Per column p of CV[n][P]: the lowest row of the minimum (the argmax of MU0 = LSM(-CV)).  Block partials.
__global__ __launch_bounds__(256) void k_col_argmin(int64_t n, int P, const float* __restrict__ CV, float* __restrict__ pv,
                                                    int* __restrict__ pi) {
  __shared__ float rv[256];
  __shared__ int ri[256];
  const int t = threadIdx.x, RP = 256 / P;
  const bool active = t < RP * P;
  const int p = active ? t % P : 0;
  float best = INFINITY;
  int bi = INT_MAX;
  if (active)
    for (int64_t v = (int64_t)blockIdx.x * RP + t / P; v < n; v += (int64_t)gridDim.x * RP) {
      const float x = CV[v * P + p];
      if (x < best) {
        best = x;
        bi = (int)v;
      }
    }
  rv[t] = best;
  ri[t] = bi;
  __syncthreads();
  if (t < P) {
    for (int r = 1; r < RP; ++r) {
      const float x = rv[r * P + t];
      const int xi = ri[r * P + t];
      if (x < best || (x == best && xi < bi)) {
        best = x;
        bi = xi;
      }
    }
    pv[(int64_t)blockIdx.x * P + t] = best;
    pi[(int64_t)blockIdx.x * P + t] = bi;
  }
}

__global__ void k_col_argmin_final(int nparts, int P, const float* __restrict__ pv, const int* __restrict__ pi,
                                   int* __restrict__ am) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float best = INFINITY;
  int bi = 0;
  for (int b = 0; b < nparts; ++b) {
    const float x = pv[(int64_t)b * P + p];
    const int xi = pi[(int64_t)b * P + p];
    if (x < best || (x == best && xi < bi)) {
      best = x;
      bi = xi;
    }
  }
  am[p] = bi;
}

// Edge-cost gradients divided by twice the cost: G[row][q] = (gCE + (cD[src] + cq) exp(-CE)) / (2 CE), 0 for
// padding rows and zero costs (d sqrt(s) / ds is unbounded at 0; such a pair contributes nothing here).
__global__ void k_bwd_edge_coef(int64_t nrows, int Q, int Qs, const int* __restrict__ beid, const int* __restrict__ bsrc,
                                const float* __restrict__ CE, const double* __restrict__ cD,
                                const double* __restrict__ cq, float* __restrict__ G) {
  const int64_t total = nrows * Qs;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / Qs;
    const int q = (int)(i - row * Qs);
    float out = 0.f;
    if (q < Q && beid[row] >= 0) {
      const double ce = (double)CE[i];
      if (ce > 0.0 && ce != kInf) {
        const double g = (double)G[i] + (cD[(int64_t)(bsrc[row] & ~kNewVertex) * Qs + q] + cq[q]) * exp(-ce);
        out = (float)(g / (2.0 * ce));
      }
    }
    G[i] = out;
  }
}

// Vertex-cost gradients divided by twice the cost (in place): G = gCV / (2 CV), 0 where CV = 0.
__global__ void k_bwd_vertex_coef(int64_t total, const float* __restrict__ CV, float* __restrict__ G) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const float c = CV[i];
    G[i] = c > 0.f ? G[i] / (2.f * c) : 0.f;
  }
}

// Outer-product sums over rows, for q < K, k < dim:  A[q][k] = sum_r G[r][q] x_r[k]^2,  B[q][k] = sum_r G[r][q]
// x_r[k],  C[q] = sum_r G[r][q],  x_r = feats[idx ? idx[r] : r] (rows with idx < 0 skipped), one block partial each.
// For the (q, k) block [q0, q0 + 16 NQT) x [k0, k0 + 16 NKT) they are GEMMs over the rows on FP64 matrix cores
// (v_mfma_f64_16x16x4f64, four rows per step).  A-operand lane
// i + 16k holds G[row k][q0 + i], B-operand lane j + 16k holds x[row k][k0 + j] (or its square), and D[4r +
// lane/16][lane%16] sits in register r.  Products of floats are exact in FP64; the sums are FP64.  The block's four
// waves are summed in order.
typedef double d4v __attribute__((ext_vector_type(4)));
template <int NQT, int NKT, bool kIdx>
__global__ __launch_bounds__(256) void k_outer_mfma(int64_t nrows, const int* __restrict__ idx,
                                                    const float* __restrict__ feats, int dim,
                                                    const float* __restrict__ G, int Gs, int K, int q0, int k0,
                                                    double* __restrict__ part) {
  __shared__ double red[3][64][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane & 15, kr = lane >> 4;
  d4v accA[NQT][NKT], accB[NQT][NKT];
#pragma unroll
  for (int a = 0; a < NQT; ++a)
#pragma unroll
    for (int b = 0; b < NKT; ++b) accA[a][b] = accB[a][b] = d4v{0.0, 0.0, 0.0, 0.0};
  double cs[NQT];
#pragma unroll
  for (int a = 0; a < NQT; ++a) cs[a] = 0.0;
  // Steps of kOS x 4 rows, software-pipelined: the raw values of the next step are loaded while the current one
  // runs on the matrix cores, and the row ids one step earlier still, so that every load of a step issues at once.
  // All loads use valid (clamped) addresses and write each register once; the values that lie past K or dim or
  // belong to no row are zeroed when they are converted, so the loads stay in flight until then.
  constexpr int kOS = 4;
  const int64_t stride = (int64_t)gridDim.x * 4 * (4 * kOS);
  float gf[kOS][NQT], xf[kOS][NKT];
  int64_t srcl[kOS];  // source rows of the loaded step (-1: none)
  int sid[kOS];       // raw ids of the step after it
  auto load_ids = [&](int64_t r0, int* sv) {
#pragma unroll
    for (int h = 0; h < kOS; ++h) {
      const int64_t row = r0 + 4 * h + kr;
      sv[h] = kIdx ? idx[row < nrows ? row : nrows - 1] : 0;
    }
  };
  auto src_of = [&](int64_t r0, int h, int raw) -> int64_t {
    const int64_t row = r0 + 4 * h + kr;
    return row < nrows ? (kIdx ? (int64_t)raw : row) : -1;
  };
  auto load = [&](int64_t r0) {
#pragma unroll
    for (int h = 0; h < kOS; ++h) {
      const int64_t row = r0 + 4 * h + kr;
      const int64_t rc = row < nrows ? row : nrows - 1, sc = srcl[h] >= 0 ? srcl[h] : 0;
#pragma unroll
      for (int a = 0; a < NQT; ++a) {
        const int q = q0 + 16 * a + i;
        gf[h][a] = G[rc * Gs + (q < K ? q : K - 1)];
      }
#pragma unroll
      for (int b = 0; b < NKT; ++b) {
        const int k = k0 + 16 * b + i;
        xf[h][b] = feats[sc * dim + (k < dim ? k : dim - 1)];
      }
    }
  };
  int64_t r0 = ((int64_t)blockIdx.x * 4 + wave) * (4 * kOS);
  if (r0 < nrows) {
    int s0[kOS];
    load_ids(r0, s0);
#pragma unroll
    for (int h = 0; h < kOS; ++h) srcl[h] = src_of(r0, h, s0[h]);
    load(r0);
    load_ids(r0 + stride, sid);
  }
  for (; r0 < nrows; r0 += stride) {
    double g[kOS][NQT], x[kOS][NKT];
#pragma unroll
    for (int h = 0; h < kOS; ++h) {
      const bool ok = srcl[h] >= 0;
#pragma unroll
      for (int a = 0; a < NQT; ++a) g[h][a] = (ok && q0 + 16 * a + i < K) ? (double)gf[h][a] : 0.0;
#pragma unroll
      for (int b = 0; b < NKT; ++b) x[h][b] = (ok && k0 + 16 * b + i < dim) ? (double)xf[h][b] : 0.0;
    }
    // The next step's loads are issued unconditionally (past the last step they read clamped, valid rows and are
    // never used), so that no branch separates them from the loop back-edge.
#pragma unroll
    for (int h = 0; h < kOS; ++h) {
      int raw = sid[h];
      if (kIdx) asm volatile("" : "+v"(raw));  // an empty register barrier: the id is first needed here
      srcl[h] = src_of(r0 + stride, h, raw);
    }
    load(r0 + stride);
    load_ids(r0 + 2 * stride, sid);
#pragma unroll
    for (int h = 0; h < kOS; ++h)
#pragma unroll
      for (int a = 0; a < NQT; ++a) {
        cs[a] += g[h][a];
#pragma unroll
        for (int b = 0; b < NKT; ++b) {
          accB[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(g[h][a], x[h][b], accB[a][b], 0, 0, 0);
          accA[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(g[h][a], x[h][b] * x[h][b], accA[a][b], 0, 0, 0);
        }
      }
  }
  // C: the four row slots of each column, then the waves.
#pragma unroll
  for (int a = 0; a < NQT; ++a) {
    cs[a] += __shfl_xor(cs[a], 16, 64);
    cs[a] += __shfl_xor(cs[a], 32, 64);
  }
  double* out = part + (int64_t)blockIdx.x * (2 * (int64_t)K * dim + K);
  auto block_sum = [&](d4v v) {
    __syncthreads();
    if (wave > 0)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[wave - 1][lane][r] = v[r];
    __syncthreads();
    if (wave == 0)
      for (int w = 0; w < 3; ++w)
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] += red[w][lane][r];
    return v;
  };
#pragma unroll
  for (int a = 0; a < NQT; ++a)
#pragma unroll
    for (int b = 0; b < NKT; ++b)
#pragma unroll
      for (int which = 0; which < 2; ++which) {
        const d4v v = block_sum(which ? accB[a][b] : accA[a][b]);
        if (wave == 0)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int q = q0 + 16 * a + 4 * r + kr, k = k0 + 16 * b + i;
            if (q < K && k < dim) out[(which ? (int64_t)K * dim : 0) + (int64_t)q * dim + k] = v[r];
          }
      }
  if (k0 == 0)
#pragma unroll
    for (int a = 0; a < NQT; ++a) {
      const d4v v = block_sum(d4v{cs[a], 0.0, 0.0, 0.0});
      const int q = q0 + 16 * a + i;
      if (wave == 0 && kr == 0 && q < K) out[2 * (int64_t)K * dim + q] = v[0];
    }
}

__global__ void k_fill_int(int* __restrict__ p, int64_t count, int value) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

// ---------------------------------------------------------------------------------------------------------------
// Launchers.

static int grid_for(int64_t total) { return (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 1 << 15)); }

void launch_fill_int(int* p, int64_t count, int value, hipStream_t st) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(count)), dim3(256), 0, st, p, count, value);
}

static int sum_blocks(int C) { return std::max(1, C); }

void launch_sum_parts(int nparts, int C, const double* part, double* out, hipStream_t st) {
  if (C == 0) return;
  hipLaunchKernelGGL(k_sum_parts, dim3(sum_blocks(C)), dim3(256), 0, st, nparts, C, part, out);
}

void launch_colsum(int64_t n, int C, const float* X, double* part, int nparts, double* out, hipStream_t st) {
  if (C == 0) return;
  hipLaunchKernelGGL(k_colsum, dim3(nparts), dim3(256), 0, st, n, C, X, part);
  hipLaunchKernelGGL(k_sum_parts, dim3(sum_blocks(C)), dim3(256), 0, st, nparts, C, part, out);
}

void launch_bwd_out(const BwdGraph& G, const double* MU, const double* colsum, float* gMU, float* gCV, hipStream_t st) {
  hipLaunchKernelGGL(k_bwd_out, dim3(grid_for(G.n * G.P)), dim3(256), 0, st, G.n * G.P, G.P, MU, colsum, gMU, gCV);
}

void launch_bwd_route(const BwdGraph& G, const BwdState& S, const float* gZ, int with_gz, const float* gVF,
                      const float* gVR, float* gEF, float* gER, unsigned* gmax, float* gERsum, double* part,
                      int nparts, double* red4, hipStream_t st) {
  if (G.Q == 0) return;
  hipLaunchKernelGGL(k_bwd_route, dim3(nparts), dim3(256), 0, st, G, S, gZ, with_gz, gVF, gVR, gEF, gER, gmax, gERsum,
                     part);
  hipLaunchKernelGGL(k_sum_parts, dim3(sum_blocks(4 * G.Q)), dim3(256), 0, st, nparts, 4 * G.Q, part, red4);
}

void launch_bwd_in(const BwdGraph& G, const BwdState& S, const double* MU, const QStat* qs1, const double* red4,
                   const float* gER, float* gVR, float* gVF, double* cD, const FmaxTerms& FE, hipStream_t st) {
  if (G.Q == 0) return;
  hipLaunchKernelGGL(k_bwd_in, dim3(grid_for(G.n * G.Q)), dim3(256), 0, st, G, S, MU, qs1, red4, gER, gVR, gVF, cD, FE);
}

void launch_bwd_fscatter(const BwdGraph& G, const float* gEF, const int* FA, const int* bsrc, int64_t brows,
                         float* gCE, hipStream_t st) {
  if (G.Q == 0) return;
  hipLaunchKernelGGL(k_bwd_fscatter, dim3(grid_for(G.n * G.Qs)), dim3(256), 0, st, G.n, G.Q, G.Qs, gEF, FA, bsrc, brows,
                     gCE);
}

// RMax edge terms, once after the last step: the minimum-cost out-edge row AO[v][q] receives -gERsum[v][q] (the sum
// over the steps of gE_R[v][q]).  Each (row, q) belongs to one (v, q).
__global__ void k_bwd_rterms(int64_t n, int Q, int Qs, const float* __restrict__ gER, const int* __restrict__ AO,
                             int64_t brows, float* __restrict__ gCE) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * Qs; i += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(i % Qs);
    if (q >= Q) continue;
    const float g = gER[i];
    if (g == 0.f) continue;
    const int row = AO[i];
    if (row >= 0 && row < brows) gCE[(int64_t)row * Qs + q] += -g;
  }
}




void launch_bwd_rterms(const BwdGraph& G, const float* gERsum, const int* AO, int64_t brows, float* gCE,
                       hipStream_t st) {
  if (G.Q == 0) return;
  const int64_t nq = G.n * G.Qs;
  hipLaunchKernelGGL(k_bwd_rterms, dim3((unsigned)std::min<int64_t>((nq + 255) / 256, 1 << 16)), dim3(256), 0, st,
                     G.n, G.Q, G.Qs, gERsum, AO, brows, gCE);
}

int bwd_vtx_tile(const BwdGraph& G) { return std::max(1, std::min(64, 8192 / std::max(1, 2 * G.Qs + G.P))); }

int bwd_vtx_blocks(const BwdGraph& G) {
  const int T = bwd_vtx_tile(G);
  return (int)std::max<int64_t>(1, std::min<int64_t>((G.n + T - 1) / T, 2048));
}

void launch_bwd_vtx(const BwdGraph& G, float* gVR, const float* gVF, unsigned long long* gVRfix, const unsigned* gmax,
                    float* gMU, double* part, hipStream_t st) {
  const int T = bwd_vtx_tile(G), nb = bwd_vtx_blocks(G);
  const size_t lds = (size_t)T * (2 * G.Qs + G.P) * sizeof(float) + (size_t)(G.P + 1 + 2 * G.Q) * sizeof(int);
  hipLaunchKernelGGL(k_bwd_vtx, dim3(nb), dim3(256), lds, st, G, gVR, gVF, gVRfix, gmax, gMU, part, T);
}

void launch_bwd_prep_fill(const BwdGraph& G, const double* red4, const int* am0, float* gMU, hipStream_t st) {
  if (G.Q == 0) return;
  hipLaunchKernelGGL(k_bwd_prep_fill, dim3(1), dim3(64), 0, st, G, red4, am0, gMU);
}

void launch_argmin_out(int64_t m, int Q, int Qs, const int2* ks, const float* CE, const float* minCEo, int* AO,
                       hipStream_t st) {
  if (m == 0 || Q == 0) return;
  hipLaunchKernelGGL(k_argmin_out, dim3(grid_for(m * Q)), dim3(256), 0, st, m, Q, Qs, ks, CE, minCEo, AO);
}

void launch_col_argmin(int64_t n, int P, const float* CV, float* pv, int* pi, int nparts, int* am, hipStream_t st) {
  hipLaunchKernelGGL(k_col_argmin, dim3(nparts), dim3(256), 0, st, n, P, CV, pv, pi);
  hipLaunchKernelGGL(k_col_argmin_final, dim3((P + 255) / 256), dim3(256), 0, st, nparts, P, pv, pi, am);
}

void launch_bwd_edge_coef(int64_t nrows, int Q, int Qs, const int* beid, const int* bsrc, const float* CE,
                          const double* cD, const double* cq, float* G, hipStream_t st) {
  if (nrows == 0) return;
  hipLaunchKernelGGL(k_bwd_edge_coef, dim3(grid_for(nrows * Qs)), dim3(256), 0, st, nrows, Q, Qs, beid, bsrc, CE, cD,
                     cq, G);
}

void launch_bwd_vertex_coef(int64_t total, const float* CV, float* G, hipStream_t st) {
  hipLaunchKernelGGL(k_bwd_vertex_coef, dim3(grid_for(total)), dim3(256), 0, st, total, CV, G);
}

void launch_outer_acc(int64_t nrows, const int* idx, const float* feats, int dim, const float* G, int Gs, int K,
                      double* part, int nparts, hipStream_t st) {
  // Matrix-core path in (q, k) blocks of 32 x 32 (each block of the grid covers every row once per launch).
  for (int q0 = 0; q0 < K; q0 += 32)
    for (int k0 = 0; k0 < dim; k0 += 32)
      if (idx)
        hipLaunchKernelGGL((k_outer_mfma<2, 2, true>), dim3(nparts), dim3(256), 0, st, nrows, idx, feats, dim, G, Gs,
                           K, q0, k0, part);
      else
        hipLaunchKernelGGL((k_outer_mfma<2, 2, false>), dim3(nparts), dim3(256), 0, st, nrows, idx, feats, dim, G,
                           Gs, K, q0, k0, part);
}

}  // namespace ac
