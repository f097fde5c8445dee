// Matcher kernels for gfx950 (CDNA4, wave64).
//
// Dataflow of one iteration (src/main.cu:160-245), reorganised so that no pattern-edge x data-edge matrix other than
// the edge costs CE is ever stored, and so that every stored value is normalised (small in magnitude near the
// matches that decide the output), with all arithmetic in FP64 registers:
//   vertex pass A  : MU = LSM(-CV + sum FMax + sum RMax) (recomputed, never stored), VF = MU[pd] - FMax,
//                    VR = MU[ps] - RMax; stores VR and VFm = VF - minCEo; column maxima and the two edge
//                    log-sum-exps, using  sum_e exp(X[src e] - CE_e) = sum_v exp(X[v] - minCEo[v] + lsO[v]).
//   edge walk B    : FMaxN[v] = max over in-edges of (VR[src] - CE) - LSE_FE   (CSC order, no atomics)
//   vertex pass C  : column log-sum-exp of MU for the next iteration's normalisation.
// RMax needs no edge pass: max over out-edges of (VF[v] - CE_e) = VF[v] - min over out-edges of CE_e, exactly.
// The log-softmax of CE and CV in the reference's prep (src/main.cu:116,118) shift each row by a constant, which
// every later log-softmax removes again, so the raw costs are used.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ac_internal.hpp"

namespace ac {

namespace {

constexpr double kInf = __builtin_huge_val();

__device__ __forceinline__ void lse_add(Lse& s, double x) {
  if (x == -kInf) return;
  if (x > s.m) {
    s.t = (s.m == -kInf) ? 0.0 : (1.0 + s.t) * (double)expf((float)(s.m - x));
    s.m = x;
  } else {
    s.t += (double)expf((float)(x - s.m));
  }
}

__device__ __forceinline__ Lse lse_merge(Lse a, Lse b) {
  if (b.m == -kInf) return a;
  if (a.m == -kInf) return b;
  if (a.m >= b.m) {
    a.t += (1.0 + b.t) * exp(b.m - a.m);
    return a;
  }
  b.t += (1.0 + a.t) * exp(a.m - b.m);
  return b;
}

__device__ __forceinline__ Lse lse_empty() { return Lse{-kInf, 0.0}; }

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// cdist (src/ac.hxx:33-56), FP64 accumulation of (a - b)^2, rounded once to FP32.

__global__ void k_cdist_vertex(int64_t n, int P, int dV, const float* __restrict__ fd,
                               const double* __restrict__ pf, float* __restrict__ CV) {
  const int64_t total = n * P;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = i / P;
    const int p = (int)(i - v * P);
    const float* a = fd + v * dV;
    const double* b = pf + (int64_t)p * dV;
    double acc = 0.0;
    for (int k = 0; k < dV; ++k) {
      double d = (double)a[k] - b[k];
      acc = fma(d, d, acc);
    }
    CV[i] = (float)sqrt(acc);
  }
}

void launch_cdist_vertex(int64_t n, int P, int dV, const float* fd, const double* pf, float* CV, hipStream_t st) {
  int64_t total = n * P;
  if (total == 0) return;
  int blocks = (int)std::min<int64_t>((total + 255) / 256, 1 << 20);
  hipLaunchKernelGGL(k_cdist_vertex, dim3(blocks), dim3(256), 0, st, n, P, dV, fd, pf, CV);
}

// One thread per CSC edge k computes the whole cost row CE[k][0..Qs) for edge eid[k], four pattern edges at a time;
// pattern features pf are [dE][Qs] doubles (zero in padding columns, whose costs are written as 0).
__global__ void k_cdist_edge(int64_t m, int Q, int Qs, int dE, const float* __restrict__ fe,
                             const int* __restrict__ eid, const double* __restrict__ pf, float* __restrict__ CE) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const float* g = fe + (int64_t)eid[k] * dE;
    float4* out = reinterpret_cast<float4*>(CE + k * Qs);
    for (int q0 = 0; q0 < Qs; q0 += 4) {
      double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
      const double* b = pf + q0;
      for (int i = 0; i < dE; ++i) {
        const double x = (double)g[i];
        const double* bi = b + (int64_t)i * Qs;
        double d0 = x - bi[0], d1 = x - bi[1], d2 = x - bi[2], d3 = x - bi[3];
        a0 = fma(d0, d0, a0);
        a1 = fma(d1, d1, a1);
        a2 = fma(d2, d2, a2);
        a3 = fma(d3, d3, a3);
      }
      float4 r;
      r.x = q0 + 0 < Q ? (float)sqrt(a0) : 0.f;
      r.y = q0 + 1 < Q ? (float)sqrt(a1) : 0.f;
      r.z = q0 + 2 < Q ? (float)sqrt(a2) : 0.f;
      r.w = q0 + 3 < Q ? (float)sqrt(a3) : 0.f;
      out[q0 >> 2] = r;
    }
  }
}

void launch_cdist_edge(int64_t m, int Q, int Qs, int dE, const float* fe, const int* eid, const double* pf,
                       float* CE, hipStream_t st) {
  if (m == 0) return;
  int blocks = (int)std::min<int64_t>((m + 255) / 256, 1 << 20);
  hipLaunchKernelGGL(k_cdist_edge, dim3(blocks), dim3(256), 0, st, m, Q, Qs, dE, fe, eid, pf, CE);
}

// ---------------------------------------------------------------------------------------------------------------
// Column log-sum-exp of -X over the n rows of X[n][P] (the prep's LSM(-CV) normaliser, src/main.cu:117).

__global__ __launch_bounds__(256) void k_col_lse_neg(int64_t n, int P, const float* __restrict__ X,
                                                     Lse* __restrict__ part) {
  __shared__ Lse red[256];
  const int t = threadIdx.x;
  const int RP = 256 / P;
  const bool active = t < RP * P;
  const int p = active ? t % P : 0;
  Lse acc = lse_empty();
  if (active)
    for (int64_t v = (int64_t)blockIdx.x * RP + t / P; v < n; v += (int64_t)gridDim.x * RP)
      lse_add(acc, -(double)X[v * P + p]);
  red[t] = acc;
  __syncthreads();
  if (t < P) {
    Lse s = lse_empty();
    for (int r = 0; r < RP; ++r) s = lse_merge(s, red[r * P + t]);
    part[(int64_t)blockIdx.x * P + t] = s;
  }
}

void launch_col_lse_neg(int64_t n, int P, const float* X, Lse* part, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_col_lse_neg, dim3(nblocks), dim3(256), 0, st, n, P, X, part);
}

// Reduce per-block partials [nparts][ncol] column by column in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_reduce_lse(int nparts, int ncol, const Lse* __restrict__ part,
                                                    Lse2* __restrict__ out) {
  __shared__ Lse red[256];
  const int c = blockIdx.x, t = threadIdx.x;
  Lse s = lse_empty();
  for (int b = t; b < nparts; b += 256) s = lse_merge(s, part[(int64_t)b * ncol + c]);
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] = lse_merge(red[t], red[t + w]);
    __syncthreads();
  }
  if (t == 0) out[c] = Lse2{red[0].m, log1p(red[0].t)};
}

void launch_reduce_lse(int nparts, int ncol, const Lse* part, Lse2* out, hipStream_t st) {
  if (ncol == 0) return;
  hipLaunchKernelGGL(k_reduce_lse, dim3(ncol), dim3(256), 0, st, nparts, ncol, part, out);
}

__global__ __launch_bounds__(256) void k_reduce_q(int nparts, int Q, const QPart* __restrict__ part,
                                                  QStat* __restrict__ out) {
  __shared__ QPart red[256];
  const int q = blockIdx.x, t = threadIdx.x;
  QPart s{-kInf, -kInf, lse_empty(), lse_empty()};
  for (int b = t; b < nparts; b += 256) {
    const QPart x = part[(int64_t)b * Q + q];
    s.vfmax = fmax(s.vfmax, x.vfmax);
    s.vrmax = fmax(s.vrmax, x.vrmax);
    s.re = lse_merge(s.re, x.re);
    s.fe = lse_merge(s.fe, x.fe);
  }
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      QPart a = red[t], b = red[t + w];
      a.vfmax = fmax(a.vfmax, b.vfmax);
      a.vrmax = fmax(a.vrmax, b.vrmax);
      a.re = lse_merge(a.re, b.re);
      a.fe = lse_merge(a.fe, b.fe);
      red[t] = a;
    }
    __syncthreads();
  }
  if (t == 0) {
    const QPart& r = red[0];
    out[q] = QStat{r.vrmax, r.vfmax, r.re.m, log1p(r.re.t), r.fe.m, log1p(r.fe.t)};
  }
}

void launch_reduce_q(int nparts, int Q, const QPart* part, QStat* out, hipStream_t st) {
  if (Q == 0) return;
  hipLaunchKernelGGL(k_reduce_q, dim3(Q), dim3(256), 0, st, nparts, Q, part, out);
}

// Scalars of the prep state (src/main.cu:116-146): MU0 = LSM(-CV), RE0 = FE0 = LSM(-CE).  The column maximum of
// MU0 sits at the minimum cost, where (x - m) = 0, so max_v MU0[v][p] = -log1p(t_p).
__global__ void k_prep_scalars(int Q, const int* __restrict__ ps, const int* __restrict__ pd,
                               const Lse2* __restrict__ cv, const Lse2* __restrict__ ce, QStat* __restrict__ qs) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  qs[q] = QStat{-cv[ps[q]].l, -cv[pd[q]].l, ce[q].m, ce[q].l, ce[q].m, ce[q].l};
}

void launch_prep_scalars(int P, int Q, const int* ps, const int* pd, const Lse2* cv, const Lse2* ce, QStat* qs,
                         hipStream_t st) {
  (void)P;
  if (Q == 0) return;
  hipLaunchKernelGGL(k_prep_scalars, dim3((Q + 255) / 256), dim3(256), 0, st, Q, ps, pd, cv, ce, qs);
}

__global__ void k_fill(float* __restrict__ p, int64_t count, float value) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

void launch_fill(float* p, int64_t count, float value, hipStream_t st) {
  if (count == 0) return;
  int blocks = (int)std::min<int64_t>((count + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, p, count, value);
}

// ---------------------------------------------------------------------------------------------------------------
// Edge walks.  A group of Qs/4 lanes walks one run of consecutive edges; lane j owns pattern edges 4j..4j+3 and
// reads them as one float4 of the edge's cost row (and of the gathered VR row).  Runs cover whole vertices except
// the pieces of very long edge lists, whose partial results go to carry slots and are combined by a fixup kernel.

int walk_blocks(int64_t nruns, int Qs) {
  const int L4 = Qs / 4;
  const int G = 64 / L4;
  const int64_t per_block = (int64_t)(kWalkBlock / 64) * G;
  return (int)((nruns + per_block - 1) / per_block);
}

// Source-side statistics of the edge costs (CSR walk): for every vertex v with out-edges and every q,
//   minCEo[v][q] = min_e CE[e][q],  lsO[v][q] = log1p(sum_{other e} exp(minCEo - CE)),  VFm[v][q] = -minCEo
// (VFm of the prep state: RE0 = -CE up to a per-q constant), and per-block partials of log sum_e exp(-CE[e][q]).
__global__ __launch_bounds__(kWalkBlock) void k_walk_src(int64_t nruns, const Run* __restrict__ runs,
                                                         const int2* __restrict__ ks, const float* __restrict__ CE,
                                                         int Q, int Qs, float* __restrict__ minCEo,
                                                         float* __restrict__ lsO, float* __restrict__ VFm,
                                                         double* __restrict__ carry, Lse* __restrict__ gpart) {
  __shared__ Lse red[kWalkBlock * 4];
  const int L4 = Qs / 4, G = 64 / L4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane / L4, j = lane - g * L4;
  const int64_t r = ((int64_t)blockIdx.x * (kWalkBlock / 64) + wave) * G + g;
  Lse glob[4] = {lse_empty(), lse_empty(), lse_empty(), lse_empty()};
  if (g < G && r < nruns) {
    const Run R = runs[r];
    const float4* CE4 = reinterpret_cast<const float4*>(CE);
    Lse s[4] = {lse_empty(), lse_empty(), lse_empty(), lse_empty()};
    int cur = ks[R.kb].y;
    auto emit = [&](int v) {
      float4 mn, ls, vf;
      mn.x = (float)(-s[0].m), mn.y = (float)(-s[1].m), mn.z = (float)(-s[2].m), mn.w = (float)(-s[3].m);
      ls.x = (float)log1p(s[0].t), ls.y = (float)log1p(s[1].t), ls.z = (float)log1p(s[2].t), ls.w = (float)log1p(s[3].t);
      vf.x = (float)s[0].m, vf.y = (float)s[1].m, vf.z = (float)s[2].m, vf.w = (float)s[3].m;
      reinterpret_cast<float4*>(minCEo)[(int64_t)v * L4 + j] = mn;
      reinterpret_cast<float4*>(lsO)[(int64_t)v * L4 + j] = ls;
      reinterpret_cast<float4*>(VFm)[(int64_t)v * L4 + j] = vf;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        glob[c] = lse_merge(glob[c], s[c]);
        s[c] = lse_empty();
      }
    };
    for (int k = R.kb; k < R.ke; ++k) {
      const int2 e = ks[k];
      if (e.y != cur) {
        emit(cur);
        cur = e.y;
      }
      const float4 c = CE4[(int64_t)e.x * L4 + j];
      lse_add(s[0], -(double)c.x);
      lse_add(s[1], -(double)c.y);
      lse_add(s[2], -(double)c.z);
      lse_add(s[3], -(double)c.w);
    }
    if (R.slot >= 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        carry[((int64_t)R.slot * Qs + 4 * j + c) * 2 + 0] = s[c].m;
        carry[((int64_t)R.slot * Qs + 4 * j + c) * 2 + 1] = s[c].t;
        glob[c] = lse_merge(glob[c], s[c]);
      }
    } else {
      emit(cur);
    }
  }
  // Block partial of the global log-sum-exp, merged in a fixed (wave, group) order per pattern edge.
#pragma unroll
  for (int c = 0; c < 4; ++c) red[threadIdx.x * 4 + c] = glob[c];
  __syncthreads();
  for (int q = threadIdx.x; q < Q; q += kWalkBlock) {
    const int jj = q >> 2, c = q & 3;
    Lse acc = lse_empty();
    for (int w = 0; w < kWalkBlock / 64; ++w)
      for (int gg = 0; gg < G; ++gg) acc = lse_merge(acc, red[(w * 64 + gg * L4 + jj) * 4 + c]);
    gpart[(int64_t)blockIdx.x * Q + q] = acc;
  }
}

void launch_walk_src(int64_t nruns, const Run* runs, const int2* ks, const float* CE, int Q, int Qs, float* minCEo,
                     float* lsO, float* VFm, double* carry, Lse* gpart, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_walk_src, dim3(nblocks), dim3(kWalkBlock), 0, st, nruns, runs, ks, CE, Q, Qs, minCEo, lsO,
                     VFm, carry, gpart);
}

__global__ void k_walk_src_fixup(int nsplit, const Split* __restrict__ splits, const double* __restrict__ carry,
                                 int Q, int Qs, float* __restrict__ minCEo, float* __restrict__ lsO,
                                 float* __restrict__ VFm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)nsplit * Qs) return;
  const int sidx = (int)(i / Qs), q = (int)(i - (int64_t)sidx * Qs);
  const Split S = splits[sidx];
  Lse s = lse_empty();
  for (int c = 0; c < S.count; ++c) {
    const int64_t o = ((int64_t)(S.first + c) * Qs + q) * 2;
    s = lse_merge(s, Lse{carry[o], carry[o + 1]});
  }
  const int64_t o = (int64_t)S.v * Qs + q;
  if (q < Q) {
    minCEo[o] = (float)(-s.m);
    lsO[o] = (float)log1p(s.t);
    VFm[o] = (float)s.m;
  } else {
    minCEo[o] = 0.f;
    lsO[o] = 0.f;
    VFm[o] = 0.f;
  }
}

void launch_walk_src_fixup(int nsplit, const Split* splits, const double* carry, int Q, int Qs, float* minCEo,
                           float* lsO, float* VFm, hipStream_t st) {
  int64_t total = (int64_t)nsplit * Qs;
  if (total == 0) return;
  hipLaunchKernelGGL(k_walk_src_fixup, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, nsplit, splits,
                     carry, Q, Qs, minCEo, lsO, VFm);
}

// Destination maxima (CSC walk), src/main.cu:220-224 with the log-softmax of src/main.cu:218 folded in:
//   FMaxN[v][q] = max over in-edges e of v of (VR[src e][q] - CE[e][q])  -  log sum_e exp(FE[q][e])
// computed exactly in FP64 (a difference of two floats is exact in double) and rounded once.  With VR == nullptr
// it is the prep's max over in-edges of -CE (src/main.cu:143-146).
template <bool kWithVR>
__global__ __launch_bounds__(kWalkBlock) void k_walk_max(int64_t nruns, const Run* __restrict__ runs,
                                                         const int2* __restrict__ sd, const float* __restrict__ CE,
                                                         const float* __restrict__ VR, int Q, int Qs,
                                                         const QStat* __restrict__ qs, float* __restrict__ FMaxN,
                                                         double* __restrict__ carry) {
  const int L4 = Qs / 4, G = 64 / L4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane / L4, j = lane - g * L4;
  const int64_t r = ((int64_t)blockIdx.x * (kWalkBlock / 64) + wave) * G + g;
  if (g >= G || r >= nruns) return;
  const Run R = runs[r];
  const float4* CE4 = reinterpret_cast<const float4*>(CE);
  const float4* VR4 = reinterpret_cast<const float4*>(VR);
  double sm[4], sl[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int q = 4 * j + c;
    sm[c] = q < Q ? qs[q].mFE : 0.0;
    sl[c] = q < Q ? qs[q].lFE : 0.0;
  }
  double mx0 = -kInf, mx1 = -kInf, mx2 = -kInf, mx3 = -kInf;
  int cur = sd[R.kb].y;
  auto emit = [&](int v) {
    float4 o;
    o.x = (float)((mx0 - sm[0]) - sl[0]);
    o.y = (float)((mx1 - sm[1]) - sl[1]);
    o.z = (float)((mx2 - sm[2]) - sl[2]);
    o.w = (float)((mx3 - sm[3]) - sl[3]);
    reinterpret_cast<float4*>(FMaxN)[(int64_t)v * L4 + j] = o;
    mx0 = mx1 = mx2 = mx3 = -kInf;
  };
  int k = R.kb;
  // Four edges per step: their ids, costs and gathered rows are all requested before any is used.
  for (; k + 4 <= R.ke; k += 4) {
    int2 e[4];
    float4 c[4], w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) e[u] = sd[k + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c[u] = CE4[(int64_t)(k + u) * L4 + j];
      w[u] = kWithVR ? VR4[(int64_t)e[u].x * L4 + j] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (e[u].y != cur) {
        emit(cur);
        cur = e[u].y;
      }
      mx0 = fmax(mx0, (double)w[u].x - (double)c[u].x);
      mx1 = fmax(mx1, (double)w[u].y - (double)c[u].y);
      mx2 = fmax(mx2, (double)w[u].z - (double)c[u].z);
      mx3 = fmax(mx3, (double)w[u].w - (double)c[u].w);
    }
  }
  for (; k < R.ke; ++k) {
    const int2 e = sd[k];
    const float4 c = CE4[(int64_t)k * L4 + j];
    const float4 w = kWithVR ? VR4[(int64_t)e.x * L4 + j] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (e.y != cur) {
      emit(cur);
      cur = e.y;
    }
    mx0 = fmax(mx0, (double)w.x - (double)c.x);
    mx1 = fmax(mx1, (double)w.y - (double)c.y);
    mx2 = fmax(mx2, (double)w.z - (double)c.z);
    mx3 = fmax(mx3, (double)w.w - (double)c.w);
  }
  if (R.slot >= 0) {
    double* o = carry + (int64_t)R.slot * Qs + 4 * j;
    o[0] = mx0;
    o[1] = mx1;
    o[2] = mx2;
    o[3] = mx3;
  } else {
    emit(cur);
  }
}

void launch_walk_max(int64_t nruns, const Run* runs, const int2* sd, const float* CE, const float* VR, int Q, int Qs,
                     const QStat* qs, float* FMaxN, double* carry, hipStream_t st) {
  if (nruns == 0) return;
  const int nb = walk_blocks(nruns, Qs);
  if (VR)
    hipLaunchKernelGGL(k_walk_max<true>, dim3(nb), dim3(kWalkBlock), 0, st, nruns, runs, sd, CE, VR, Q, Qs, qs,
                       FMaxN, carry);
  else
    hipLaunchKernelGGL(k_walk_max<false>, dim3(nb), dim3(kWalkBlock), 0, st, nruns, runs, sd, CE, VR, Q, Qs, qs,
                       FMaxN, carry);
}

__global__ void k_walk_max_fixup(int nsplit, const Split* __restrict__ splits, const double* __restrict__ carry,
                                 int Q, int Qs, const QStat* __restrict__ qs, float* __restrict__ FMaxN) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)nsplit * Qs) return;
  const int sidx = (int)(i / Qs), q = (int)(i - (int64_t)sidx * Qs);
  const Split S = splits[sidx];
  double mx = -kInf;
  for (int c = 0; c < S.count; ++c) mx = fmax(mx, carry[(int64_t)(S.first + c) * Qs + q]);
  FMaxN[(int64_t)S.v * Qs + q] = q < Q ? (float)((mx - qs[q].mFE) - qs[q].lFE) : 0.f;
}

void launch_walk_max_fixup(int nsplit, const Split* splits, const double* carry, int Q, int Qs, const QStat* qs,
                           float* FMaxN, hipStream_t st) {
  int64_t total = (int64_t)nsplit * Qs;
  if (total == 0) return;
  hipLaunchKernelGGL(k_walk_max_fixup, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, nsplit, splits,
                     carry, Q, Qs, qs, FMaxN);
}

// ---------------------------------------------------------------------------------------------------------------
// Vertex passes.  A block walks tiles of VT consecutive vertices in three phases:
//   1. threads mapped to (q, row): FMax = max(fillF, FMaxN), RMax = max(fillR, VFm - LSE_RE) into LDS (FP64);
//   2. threads mapped to (p, row): MU_pre = -CV + sum of the FMax / RMax terms of p in ascending q (FMax before
//      RMax, src/ac.hxx:164-170 made deterministic); MU = MU_pre - LSE_MU;
//   3. (mode A) threads mapped to (q, row): VF = MU[pd] - FMax, VR = MU[ps] - RMax (src/main.cu:165-176).
// Every global access is row-contiguous in the thread index, so loads and stores coalesce.

int vertex_tile(int P, int Q) {
  int vt = 65536 / ((2 * Q + P) * 8);
  return vt < 1 ? 1 : (vt > 64 ? 64 : vt);
}

template <int kMode>
__global__ __launch_bounds__(kVertexBlock) void k_vertex(VertexArgs a) {
  extern __shared__ double lds[];
  const int Q = a.Q, P = a.P, Qs = a.Qs, VT = a.VT;
  double* Fm = lds;              // [VT][Q]
  double* Rm = Fm + VT * Q;      // [VT][Q]
  double* MUs = Rm + VT * Q;     // [VT][P]
  const int t = threadIdx.x;
  const int RQ = Q > 0 ? kVertexBlock / Q : 0;
  const int RP = kVertexBlock / P;
  const bool qthread = t < RQ * Q;
  const int q = qthread ? t % Q : 0, rq = qthread ? t / Q : 0;
  const bool pthread = t < RP * P;
  const int p = pthread ? t % P : 0, rp = pthread ? t / P : 0;
  const bool sums = !a.first;

  QStat s{};
  int q_ps = 0, q_pd = 0;
  if (qthread) {
    s = a.qs[q];
    q_ps = a.ps[q];
    q_pd = a.pd[q];
  }
  Lse2 ms{0.0, 0.0};
  int ib = 0, ie = 0;
  if (pthread) {
    if (kMode != kModeC) ms = a.pst[p];
    ib = a.inc_off[p];
    ie = a.inc_off[p + 1];
  }
  double vfmax = -kInf, vrmax = -kInf;
  Lse lre = lse_empty(), lfe = lse_empty(), lmu = lse_empty();

  const int64_t ntiles = (a.n + VT - 1) / VT;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t v0 = tile * VT;
    const int nv = (int)((a.n - v0) < VT ? (a.n - v0) : VT);
    if (qthread) {
      for (int r = rq; r < nv; r += RQ) {
        const int64_t i = (v0 + r) * Qs + q;
        const double fm = fmax(s.fillF, (double)a.FMaxN[i]);
        const float vfm = a.VFm[i];
        const double rn = vfm == -INFINITY ? -kInf : ((double)vfm - s.mRE) - s.lRE;
        Fm[r * Q + q] = fm;
        Rm[r * Q + q] = fmax(s.fillR, rn);
      }
    }
    __syncthreads();
    if (pthread) {
      for (int r = rp; r < nv; r += RP) {
        double acc = -(double)a.CV[(v0 + r) * P + p];
        if (sums)
          for (int x = ib; x < ie; ++x) {
            const int code = a.inc[x];
            acc += (code & 1) ? Rm[r * Q + (code >> 1)] : Fm[r * Q + (code >> 1)];
          }
        if (kMode == kModeC) {
          lse_add(lmu, acc);
        } else {
          const double mu = (acc - ms.m) - ms.l;
          if (kMode == kModeA)
            MUs[r * P + p] = mu;
          else
            a.out[(v0 + r) * P + p] = (float)mu;
        }
      }
    }
    if (kMode == kModeA) {
      __syncthreads();
      if (qthread) {
        for (int r = rq; r < nv; r += RQ) {
          const int64_t i = (v0 + r) * Qs + q;
          const double vf = MUs[r * P + q_pd] - Fm[r * Q + q];
          const double vr = MUs[r * P + q_ps] - Rm[r * Q + q];
          a.VR[i] = (float)vr;
          vfmax = fmax(vfmax, vf);
          vrmax = fmax(vrmax, vr);
          const float mo = a.minCEo[i];
          if (mo != INFINITY) {
            const double lo = (double)a.lsO[i];
            const double x = vf - (double)mo;
            a.VFm[i] = (float)x;
            lse_add(lre, x + lo);
            lse_add(lfe, (vr - (double)mo) + lo);
          } else {
            a.VFm[i] = -INFINITY;
          }
        }
      }
    }
    __syncthreads();
  }

  // Block partials, merged over the rows of each column in a fixed order.
  if (kMode == kModeA) {
    QPart* red = reinterpret_cast<QPart*>(lds);
    if (qthread) red[rq * Q + q] = QPart{vfmax, vrmax, lre, lfe};
    __syncthreads();
    if (t < Q) {
      QPart acc{-kInf, -kInf, lse_empty(), lse_empty()};
      for (int r = 0; r < RQ; ++r) {
        const QPart x = red[r * Q + t];
        acc.vfmax = fmax(acc.vfmax, x.vfmax);
        acc.vrmax = fmax(acc.vrmax, x.vrmax);
        acc.re = lse_merge(acc.re, x.re);
        acc.fe = lse_merge(acc.fe, x.fe);
      }
      a.qpart[(int64_t)blockIdx.x * Q + t] = acc;
    }
  } else if (kMode == kModeC) {
    Lse* red = reinterpret_cast<Lse*>(lds);
    if (pthread) red[rp * P + p] = lmu;
    __syncthreads();
    if (t < P) {
      Lse acc = lse_empty();
      for (int r = 0; r < RP; ++r) acc = lse_merge(acc, red[r * P + t]);
      a.ppart[(int64_t)blockIdx.x * P + t] = acc;
    }
  }
}

void launch_vertex(int mode, const VertexArgs& a, int nblocks, hipStream_t st) {
  size_t tiles = (size_t)a.VT * (2 * a.Q + a.P) * sizeof(double);
  size_t red = std::max(sizeof(QPart) * kVertexBlock, sizeof(Lse) * kVertexBlock);
  size_t lds = std::max(tiles, red);
  if (mode == kModeA)
    hipLaunchKernelGGL(k_vertex<kModeA>, dim3(nblocks), dim3(kVertexBlock), lds, st, a);
  else if (mode == kModeC)
    hipLaunchKernelGGL(k_vertex<kModeC>, dim3(nblocks), dim3(kVertexBlock), lds, st, a);
  else
    hipLaunchKernelGGL(k_vertex<kModeOut>, dim3(nblocks), dim3(kVertexBlock), lds, st, a);
}

}  // namespace ac
