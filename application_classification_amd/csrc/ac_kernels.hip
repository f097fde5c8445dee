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

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "ac_internal.hpp"
#include "ac_lse.hpp"
#include "ac_sqrt_rn.hpp"

namespace ac {

namespace {

constexpr double kInf = __builtin_huge_val();

// Edge costs are row-major [rows][Qs].
__device__ __forceinline__ int64_t ce_index(int64_t k, int q, int Qs) { return k * Qs + q; }

// Native 4-wide vector (register-resident across loop iterations, unlike the HIP float4 wrapper struct).
typedef float f4v __attribute__((ext_vector_type(4)));

// Streaming loads (__builtin_nontemporal_load): data read once per pass (edge costs, edge ids) is loaded
// non-temporally so that it does not evict the gathered rows that are reused from L2.

// Packed costs: four 24-bit codes in three dwords, c0 | c1 << 24, c1 >> 8 | c2 << 16, c2 >> 16 | c3 << 8.  A code
// is the cost's bit pattern minus the bundle's base (a biased exponent << 23, so one bundle spans two binades);
// code 0 stands for +inf (padding rows and columns past Q).
struct U3 {
  unsigned x, y, z;
};
__device__ __forceinline__ U3 nt_load_u3(const unsigned* p) {
  return U3{__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + 1), __builtin_nontemporal_load(p + 2)};
}
__device__ __forceinline__ float decode_cost(unsigned c, unsigned base) {
  return c ? __uint_as_float(base + c) : INFINITY;
}
__device__ __forceinline__ f4v decode_costs(U3 w, unsigned base) {
  const unsigned c0 = w.x & 0xFFFFFFu;
  const unsigned c1 = __builtin_amdgcn_alignbit(w.y, w.x, 24) & 0xFFFFFFu;
  const unsigned c2 = __builtin_amdgcn_alignbit(w.z, w.y, 16) & 0xFFFFFFu;
  const unsigned c3 = w.z >> 8;
  return f4v{decode_cost(c0, base), decode_cost(c1, base), decode_cost(c2, base), decode_cost(c3, base)};
}

// max with the position of the maximum: the larger value wins, equal values keep the lower vertex (-1 = none).
__device__ __forceinline__ void argmax_merge(double& m, int& a, double x, int b) {
  if (x > m || (x == m && b >= 0 && (a < 0 || b < a))) {
    m = x;
    a = b;
  }
}

__device__ __forceinline__ void qpart_merge(QPart& a, const QPart& b) {
  argmax_merge(a.vfmax, a.avf, b.vfmax, b.avf);
  argmax_merge(a.vrmax, a.avr, b.vrmax, b.avr);
  a.re = lse_merge(a.re, b.re);
  a.fe = lse_merge(a.fe, b.fe);
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// cdist (src/ac.hxx:33-56): out[r][c] = || feats[row r] - pat_c ||, FP64 accumulation of (a - b)^2 rounded once to
// FP32.  Pattern features are [Kp / 8][dim][8] doubles (Kp = K rounded up to 8, zero padded: each group of eight
// patterns is one contiguous block, eight values per feature); row r of the output reads
// feature row idx[r] (idx == nullptr: row r).  Columns K <= c < Kstore of the output are written as 0, and rows with
// idx[r] < 0 (padding of the bundle layout) as +inf.  With w != nullptr the distance is weighted per dimension,
// sqrt(sum_k w_k (a_k - b_k)^2) (the learnable feature metric of the adjoint).

constexpr int kCdTile = 128;

// A tile of kCdTile feature rows is staged in LDS with row-contiguous (coalesced) loads; each thread then computes
// one row's distances to eight patterns at a time, the pattern values being wave-uniform (scalar loads).
template <bool kW>
__global__ __launch_bounds__(kCdTile) void k_cdist_rows(int64_t nrows, const int* __restrict__ idx,
                                                        const float* __restrict__ feats, int dim,
                                                        const double* __restrict__ pat, int K, int Kp, int Kstore,
                                                        float* __restrict__ out, int out_stride, int tile,
                                                        const double* __restrict__ w) {
  extern __shared__ float rows[];  // [tile][dim + 1], tile <= kCdTile
  const int ld = dim + 1;
  for (int64_t base = blockIdx.x * (int64_t)tile; base < nrows; base += (int64_t)gridDim.x * tile) {
    const int nt = (int)((nrows - base) < tile ? (nrows - base) : tile);
    __syncthreads();
    for (int e = threadIdx.x; e < nt * dim; e += kCdTile) {
      const int r = e / dim, c = e - r * dim;
      const int64_t src = idx ? (int64_t)idx[base + r] : base + r;
      rows[r * ld + c] = src >= 0 ? feats[src * dim + c] : 0.f;
    }
    __syncthreads();
    if ((int)threadIdx.x < nt) {
      const float* x = rows + threadIdx.x * ld;
      const int64_t row = base + threadIdx.x;
      const bool pad = idx && idx[row] < 0;
      float* o = out + row * out_stride;
      for (int k0 = 0; k0 < Kstore; k0 += 8) {
        double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4
        for (int i = 0; i < dim; ++i) {
          const double xi = (double)x[i];
          const double* pv = pat + ((int64_t)(k0 / 8) * dim + i) * 8;
          const double wi = kW ? w[i] : 1.0;
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const double d = xi - pv[c];
            a[c] = fma(kW ? wi * d : d, d, a[c]);
          }
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int k = k0 + c;
          if (k < Kstore) o[k] = pad ? INFINITY : k < K ? (float)sqrt(a[c]) : 0.f;
        }
      }
    }
  }
}

// FP64 matrix cores (v_mfma_f64_16x16x4f64), GEMM form: d^2 = |x|^2 + |p|^2 - 2 x.p with the weights folded in
// (x.p -> sum w x p, |x|^2 -> sum w x^2, |p|^2 -> sum w p^2).  One wave takes 16 feature rows at a time and forms
// D = C + A B with the patterns as A (16 pattern columns by 4 features per instruction, scaled by -2) and the rows as
// B (4 features by 16 rows), accumulating from C = |x|^2 + |p|^2, so D is d^2 itself.  The operand k index is
// permuted so that each lane's B fragment is one contiguous chunk of its row: lane l = j + 16 kk holds row j's
// features [kk D4, (kk + 1) D4), and step s of the product takes feature kk D4 + s from both operands (the sum over k
// is order-free in exact arithmetic and the accumulation is FP64).  Register r of lane l holds D[4 r + l / 16][l % 16]:
// A row m = 4 r + g is staged as pattern column 4 g + r of its tile, so lane j + 16 g ends with four consecutive
// columns 4 g .. 4 g + 3 of its own row j.  Its |x|^2, row id and output address are then its own (no lane
// exchange), and the four distances leave as one 16-byte store.  Pairs with d^2 below 2^-10 (|x|^2 + |p|^2) (with
// the largest |p|^2 of the lane's four columns, a superset) lose digits to cancellation and are recomputed directly
// as sum w (x - p)^2 in FP64, so the result keeps the direct form's accuracy (relative error of d^2 below 1e-11
// elsewhere).
typedef double d4v __attribute__((ext_vector_type(4)));

// Pattern column of MFMA row m within a tile (and the reverse: the map is its own inverse).
__device__ __forceinline__ int mfma_col(int m) { return 4 * (m & 3) + (m >> 2); }

// LDS position of pattern operand e = c * 4 D4 + d (pattern column c = 16 pt + mfma_col(m), feature d = D4 kk + s2):
// the layout [pt][s2][kk][m], so that product step s2 of tile pt reads 64 consecutive doubles across the wave (lane
// 16 kk + m), one conflict-free LDS access.  In the layout of patm itself the 64 lanes' operands lie 128 or 512
// bytes apart, all in one bank pair, and each access is serialised.
__device__ __forceinline__ int lds_operand(int e, int D4) {
  const int c = e / (4 * D4), d = e - c * (4 * D4);
  const int pt = c >> 4, m = mfma_col(c & 15), kk = d / D4, s2 = d - kk * D4;
  return ((pt * D4 + s2) * 4 + kk) * 16 + m;
}

// Correctly rounded square root of x >= 0 in FP32 (x >= 2^-104): the hardware root corrected by sqrt_rn_fix, eight
// VALU instructions against the library's general sequence.
__device__ __forceinline__ float sqrt_rn(float x) { return sqrt_rn_fix(x, __builtin_amdgcn_sqrtf(x)); }

// Up to 64 features the kernel fits three waves per SIMD (at most 168 registers), which keeps the matrix core busy
// while other waves run their epilogues and loads; wider rows keep the whole register file.  kW: weighted metric (a
// template parameter, so the unweighted loop carries no per-feature select).  kVec: the row width is exactly 4 D4
// (e.g. 32 or 64 features), so each lane's 16-float chunk loads as float4s with no feature past dim; otherwise each
// feature loads on its own at a clamped index (a separate instantiation, whose per-lane offsets would otherwise
// occupy registers in the float4 loop too).
template <int D4, bool kIdx, bool kW, bool kVec>
__global__ __launch_bounds__(256, D4 <= 16 ? 3 : 1) void k_cdist_mfma(int64_t nrows, const int* __restrict__ idx,
                                                    const float* __restrict__ feats, int dim,
                                                    const double* __restrict__ patm, const double* __restrict__ pp,
                                                    const double* __restrict__ w, const double* __restrict__ pat,
                                                    int K, int Kstore, float* __restrict__ out, int out_stride) {
  // LDS: npt * 16 * 4 D4 pattern operands (the block's copy of patm, laid out by lds_operand), [4 D4] weights,
  // [npt * 16] pattern norms.
  extern __shared__ double bsh[];
  const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
  const int64_t nw = (int64_t)gridDim.x * 4, step = nw * 16;
  const int npt = (Kstore + 15) / 16;
  double* wsh = bsh + npt * 16 * 4 * D4;
  double* psh = wsh + 4 * D4;
  for (int e = threadIdx.x; e < npt * 16 * 4 * D4; e += 256) bsh[lds_operand(e, D4)] = patm[e];
  if (kW)
    for (int d = threadIdx.x; d < 4 * D4; d += 256) wsh[d] = d < dim ? w[d] : 0.0;
  for (int c = threadIdx.x; c < npt * 16; c += 256) psh[c] = pp[c];
  __syncthreads();
  // Without kVec the features past dim are zeroed at conversion (lim of this lane's chunk exist).  Rows that do not
  // exist are not zeroed: their loads read row 0 and their outputs are +inf regardless.
  static_assert(D4 % 4 == 0, "chunks of whole float4s");
  constexpr bool vec = kVec;
  const int lim = dim - kk * D4;
  // Four columns leave as one 16-byte store when every group of four starts a 16-byte aligned address.
  const bool vst = (Kstore & 3) == 0 && (out_stride & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  // Feature row of lane row i in the tile at r (-1: past the end or a padding row), in two steps so that the id load
  // stays in flight: row_raw loads idx at a valid (clamped) row, row_of interprets the loaded value later.
  auto row_raw = [&](int64_t r) -> int {
    const int64_t row = r + i;
    return kIdx ? idx[row < nrows ? row : nrows - 1] : 0;
  };
  auto row_of = [&](int64_t r, int raw) -> int64_t {
    const int64_t row = r + i;
    return row < nrows ? (kIdx ? (int64_t)raw : row) : -1;
  };
  // This lane's chunk [kk D4, (kk + 1) D4) of row src as raw floats.  Every load has a valid address (row 0 for
  // src < 0, the last feature past dim) and no lane's register is written twice, so the loads stay in flight until
  // the values are converted.
  auto load_x = [&](int64_t src, float* x) {
    const float* xs = feats + (src >= 0 ? src : 0) * dim;
    if constexpr (vec) {
      const f4v* x4 = reinterpret_cast<const f4v*>(xs + kk * D4);
#pragma unroll
      for (int u = 0; u < D4 / 4; ++u) {
        const f4v v = x4[u];
#pragma unroll
        for (int c = 0; c < 4; ++c) x[4 * u + c] = v[c];
      }
    } else {
#pragma unroll
      for (int s2 = 0; s2 < D4; ++s2) {
        const int d = kk * D4 + s2;
        x[s2] = xs[d < dim ? d : dim - 1];
      }
    }
  };
  auto convert = [&](const float* x, double* a) {
    if constexpr (vec) {
#pragma unroll
      for (int s2 = 0; s2 < D4; ++s2) a[s2] = (double)x[s2];
    } else {
#pragma unroll
      for (int s2 = 0; s2 < D4; ++s2) a[s2] = s2 < lim ? (double)x[s2] : 0.0;
    }
  };
  // d^2 of feature row src against pattern column c, directly in FP64 (the cancellation fallback).
  auto direct = [&](int64_t src, int c) -> double {
    const float* xs = feats + src * dim;
    double d2 = 0.0;
    for (int d = 0; d < dim; ++d) {
      const double e = (double)xs[d] - pat[((int64_t)(c / 8) * dim + d) * 8 + c % 8];
      d2 = fma(kW ? w[d] * e : e, e, d2);
    }
    return d2;
  };
  // The next tile's features are in flight (raw floats) during this tile's products and are converted at the end of
  // the iteration; its row ids are loaded a tile earlier still, so that its gather issues at once.
  float xr[D4];
  double a[D4];
  int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  int64_t src = -1, srcn = -1;
  if (r0 < nrows) {
    src = row_of(r0, row_raw(r0));
    srcn = row_of(r0 + step, row_raw(r0 + step));
    load_x(src, xr);
    convert(xr, a);
  }
  for (; r0 < nrows; r0 += step) {
    const bool more = r0 + step < nrows;
    if (more) load_x(srcn, xr);
    const int rawnn = row_raw(r0 + 2 * step);
    double xx = 0.0;
#pragma unroll
    for (int s2 = 0; s2 < D4; ++s2) xx = fma(kW ? wsh[kk * D4 + s2] * a[s2] : a[s2], a[s2], xx);
    xx += __shfl_xor(xx, 16, 64);
    xx += __shfl_xor(xx, 32, 64);  // every lane: |x|^2 of its row i
    const bool live = src >= 0, rok = r0 + i < nrows;
    float* orow = out + (rok ? r0 + i : 0) * out_stride;
    // C = |x|^2 + |p|^2 of this lane's four columns 16 t + 4 kk + r of tile t, and the largest |p|^2 among them.
    auto init = [&](int t, d4v& acc, double& pm) {
      const double* pc = psh + t * 16 + 4 * kk;
      pm = fmax(fmax(pc[0], pc[1]), fmax(pc[2], pc[3]));
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = xx + pc[r];
    };
    // Epilogue of tile t, branch-free but for the rare cancellation: the square root in FP32 (sqrt_rn of the rounded
    // d^2, within 0.75 ulp of the distance, since FP64 VALU work here competes with the products for the same FP64
    // rate), +inf for rows that do not exist, 0 in the padding columns [K, Kstore).
    auto emit = [&](int t, const d4v& acc, double pm) {
      const int cb = t * 16 + 4 * kk;
      const double thr = 0x1p-10 * (xx + pm);
      double d2[4];
      bool fix = false;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        d2[r] = acc[r];
        fix |= live && cb + r < K && d2[r] < thr;
      }
      if (__any(fix)) {  // cancellation: recompute those pairs directly
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (live && cb + r < K && d2[r] < thr) d2[r] = direct(src, cb + r);
      }
      // The root is taken unconditionally (of 0 in the padding columns and rows that do not exist), so no lane
      // branches around it.
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float s = sqrt_rn(live && cb + r < K ? (float)d2[r] : 0.f);
        v[r] = live ? s : INFINITY;
      }
      if (vst) {
        if (rok && cb < Kstore) *reinterpret_cast<f4v*>(orow + cb) = f4v{v[0], v[1], v[2], v[3]};
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (rok && cb + r < Kstore) orow[cb + r] = v[r];
      }
    };
    // Pattern tiles in pairs: two independent accumulator chains interleave on the matrix core.
    for (int pt = 0; pt < npt; pt += 2) {
      const double* bp0 = bsh + (pt * D4 * 4 + kk) * 16 + i;
      d4v acc0, acc1;
      double pm0, pm1;
      init(pt, acc0, pm0);
      if (pt + 1 < npt) {
        const double* bp1 = bp0 + D4 * 64;
        init(pt + 1, acc1, pm1);
#pragma unroll
        for (int s2 = 0; s2 < D4; ++s2) {
          acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(bp0[s2 * 64], a[s2], acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(bp1[s2 * 64], a[s2], acc1, 0, 0, 0);
        }
        emit(pt, acc0, pm0);
        emit(pt + 1, acc1, pm1);
      } else {
#pragma unroll
        for (int s2 = 0; s2 < D4; ++s2) acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(bp0[s2 * 64], a[s2], acc0, 0, 0, 0);
        emit(pt, acc0, pm0);
      }
    }
    if (more) convert(xr, a);
    // An empty register barrier: the loaded id is first read here, so its load stays in flight until now.
    int raw = rawnn;
    if (kIdx) asm volatile("" : "+v"(raw));
    src = srcn;
    srcn = row_of(r0 + 2 * step, raw);
  }
}

// MFMA operands of the patterns: patm[c][4 D4] = -2 w_d p_c[d] (zero past dim or past K), pp[c] = sum_d w_d p_c[d]^2.
// The factor -2 is exact.
__global__ void k_build_patm(int K, int Kp16, int dim, int D4, const double* __restrict__ pat,
                             const double* __restrict__ w, double* __restrict__ patm, double* __restrict__ pp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Kp16) return;
  double s2 = 0.0;
  for (int d = 0; d < 4 * D4; ++d) {
    double v = 0.0;
    if (c < K && d < dim) {
      const double pv = pat[((int64_t)(c / 8) * dim + d) * 8 + c % 8];
      const double wd = w ? w[d] : 1.0;
      v = wd * pv;
      s2 += v * pv;
    }
    patm[(int64_t)c * (4 * D4) + d] = -2.0 * v;
  }
  pp[c] = s2;
}

int cdist_mfma_d4(int dim) {
  if (dim <= 16) return 4;
  if (dim <= 32) return 8;
  if (dim <= 64) return 16;
  if (dim <= 128) return 32;
  return 0;
}

size_t cdist_scratch_doubles(int K, int dim) {
  const int D4 = cdist_mfma_d4(dim);
  const int Kp16 = (K + 15) / 16 * 16;
  return D4 ? (size_t)Kp16 * (4 * D4 + 1) : 0;
}

// Fallback for feature widths too large to stage: one thread per output element, rows read from global memory.
__global__ void k_cdist_direct(int64_t nrows, const int* __restrict__ idx, const float* __restrict__ feats, int dim,
                               const double* __restrict__ pat, int K, int Kp, int Kstore, float* __restrict__ out,
                               int out_stride, const double* __restrict__ w) {
  const int64_t total = nrows * Kstore;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / Kstore;
    const int k = (int)(e - r * Kstore);
    const int64_t src = idx ? (int64_t)idx[r] : r;
    const float* x = feats + src * dim;
    double a = 0.0;
    if (k < K && src >= 0)
      for (int i = 0; i < dim; ++i) {
        const double d = (double)x[i] - pat[((int64_t)(k / 8) * dim + i) * 8 + k % 8];
        a = fma(w ? w[i] * d : d, d, a);
      }
    out[r * out_stride + k] = src < 0 ? INFINITY : k < K ? (float)sqrt(a) : 0.f;
  }
}

void launch_cdist(int64_t nrows, const int* idx, const float* feats, int dim, const double* pat, int K, int Kp,
                  int Kstore, float* out, int out_stride, const double* w, double* scratch, hipStream_t st) {
  if (nrows == 0 || Kstore == 0) return;
  const int D4 = cdist_mfma_d4(dim);
  // The pattern operands, weights and norms are staged in LDS: up to 64 KB.
  if (scratch && D4 && dim > 0 &&
      ((size_t)((Kstore + 15) / 16 * 16 + 1) * 4 * D4 + (Kstore + 15) / 16 * 16) * sizeof(double) <= 65536) {
    const int Kp16 = (Kstore + 15) / 16 * 16;
    double* patm = scratch;
    double* pp = scratch + (size_t)Kp16 * 4 * D4;
    hipLaunchKernelGGL(k_build_patm, dim3((Kp16 + 63) / 64), dim3(64), 0, st, K, Kp16, dim, D4, pat, w, patm, pp);
    const int blocks = (int)std::min<int64_t>((nrows + 63) / 64, 8192);
    const size_t lds = ((size_t)(Kp16 + 1) * 4 * D4 + Kp16) * sizeof(double);  // operands, weights, norms
#define AC_CD_MFMA_V(N, V)                                                                                        \
  if (idx && w)                                                                                                   \
    hipLaunchKernelGGL((k_cdist_mfma<N, true, true, V>), dim3(blocks), dim3(256), lds, st, nrows, idx, feats, dim, \
                       patm, pp, w, pat, K, Kstore, out, out_stride);                                             \
  else if (idx)                                                                                                   \
    hipLaunchKernelGGL((k_cdist_mfma<N, true, false, V>), dim3(blocks), dim3(256), lds, st, nrows, idx, feats,     \
                       dim, patm, pp, w, pat, K, Kstore, out, out_stride);                                        \
  else if (w)                                                                                                     \
    hipLaunchKernelGGL((k_cdist_mfma<N, false, true, V>), dim3(blocks), dim3(256), lds, st, nrows, idx, feats,     \
                       dim, patm, pp, w, pat, K, Kstore, out, out_stride);                                        \
  else                                                                                                            \
    hipLaunchKernelGGL((k_cdist_mfma<N, false, false, V>), dim3(blocks), dim3(256), lds, st, nrows, idx, feats,    \
                       dim, patm, pp, w, pat, K, Kstore, out, out_stride);
#define AC_CD_MFMA(N)                \
  case N:                            \
    if (dim == 4 * N) {              \
      AC_CD_MFMA_V(N, true)          \
    } else {                         \
      AC_CD_MFMA_V(N, false)         \
    }                                \
    return;
    switch (D4) {
      AC_CD_MFMA(4)
      AC_CD_MFMA(8)
      AC_CD_MFMA(16)
      AC_CD_MFMA(32)
      default:
        break;
    }
#undef AC_CD_MFMA
#undef AC_CD_MFMA_V
  }
  if (dim <= 255) {
    // At most 64 KB of LDS: 128 rows up to 127 features, fewer rows beyond.
    const int tile = std::min(kCdTile, 16384 / (dim + 1));
    const size_t lds = (size_t)tile * (dim + 1) * sizeof(float);
    const int blocks = (int)std::min<int64_t>((nrows + tile - 1) / tile, 1 << 16);
    if (w)
      hipLaunchKernelGGL(k_cdist_rows<true>, dim3(blocks), dim3(kCdTile), lds, st, nrows, idx, feats, dim, pat, K,
                         Kp, Kstore, out, out_stride, tile, w);
    else
      hipLaunchKernelGGL(k_cdist_rows<false>, dim3(blocks), dim3(kCdTile), lds, st, nrows, idx, feats, dim, pat, K,
                         Kp, Kstore, out, out_stride, tile, w);
  } else {
    const int blocks = (int)std::min<int64_t>((nrows * Kstore + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_cdist_direct, dim3(blocks), dim3(256), 0, st, nrows, idx, feats, dim, pat, K, Kp, Kstore,
                       out, out_stride, w);
  }
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

// Partials a thread of a fixed-order reduction loads before merging them: one memory round trip per batch rather
// than one per partial (and per field).  Slots past the last partial load the last partial (an in-bounds address)
// and then hold the empty partial, whose merge is an exact identity, so the merge sequence is the same for every
// batch size.
constexpr int kRedBatchLse = 8, kRedBatchQ = 4;

// Reduce per-block partials [nparts][ncol] column by column in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_reduce_lse(int nparts, int ncol, const Lse* __restrict__ part,
                                                    Lse2* __restrict__ out) {
  __shared__ Lse red[256];
  const int c = blockIdx.x, t = threadIdx.x;
  Lse s = lse_empty();
  for (int b0 = t; b0 < nparts; b0 += 256 * kRedBatchLse) {
    Lse v[kRedBatchLse];
#pragma unroll
    for (int u = 0; u < kRedBatchLse; ++u) {
      const int b = b0 + u * 256;
      v[u] = part[(int64_t)(b < nparts ? b : nparts - 1) * ncol + c];
    }
#pragma unroll
    for (int u = 0; u < kRedBatchLse; ++u)
      if (b0 + u * 256 >= nparts) v[u].m = -kInf;
#pragma unroll
    for (int u = 0; u < kRedBatchLse; ++u) s = lse_merge(s, v[u]);
  }
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

// cnull != nullptr: the iteration's fills are lowered by the PNNL Cnull term, fillF = max VR - cnull[q] and
// fillR = max VF - cnull[q] (test.py:148-149; cnull <= 0, so the fills rise and the pruning bound below, which assumes
// the unraised fill, stays valid).
__global__ __launch_bounds__(256) void k_reduce_q(int nparts, int Q, const QPart* __restrict__ part,
                                                  const double* __restrict__ cnull, QStat* __restrict__ out) {
  __shared__ QPart red[256];
  const int q = blockIdx.x, t = threadIdx.x;
  const QPart empty{-kInf, -kInf, lse_empty(), lse_empty(), -1, -1};
  QPart s = empty;
  for (int b0 = t; b0 < nparts; b0 += 256 * kRedBatchQ) {
    QPart v[kRedBatchQ];
#pragma unroll
    for (int u = 0; u < kRedBatchQ; ++u) {
      const int b = b0 + u * 256;
      v[u] = part[(int64_t)(b < nparts ? b : nparts - 1) * Q + q];
    }
#pragma unroll
    for (int u = 0; u < kRedBatchQ; ++u)
      if (b0 + u * 256 >= nparts) v[u] = empty;
#pragma unroll
    for (int u = 0; u < kRedBatchQ; ++u) qpart_merge(s, v[u]);
  }
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      QPart a = red[t];
      qpart_merge(a, red[t + w]);
      red[t] = a;
    }
    __syncthreads();
  }
  if (t == 0) {
    const QPart r = red[0];
    // Edge pruning bound.  An edge term is (VR[src] - CE) - LSE_FE with VR[src] <= fillF (rounded to FP32 once),
    // so it cannot exceed fillF when CE + LSE_FE >= 2^-22 (|fillF| + |CE| + |LSE_FE|); cthr = -LSE_FE plus four
    // times that margin (a cost near the bound has |CE| ~ |LSE_FE|; a larger one clears it by itself).
    const double fF = r.vrmax, sl = r.fe.m + log1p(r.fe.t);
    double cthr = kInf;
    if (fF > -kInf && fF < kInf && sl > -kInf && sl < kInf)
      cthr = -sl + 0x1p-20 * (fabs(fF) + 2.0 * fabs(sl) + 1.0);
    const double fillF = cnull ? r.vrmax - cnull[q] : r.vrmax, fillR = cnull ? r.vfmax - cnull[q] : r.vfmax;
    out[q] = QStat{fillF, fillR, r.re.m, log1p(r.re.t), r.fe.m, log1p(r.fe.t), r.avr, r.avf, cthr};
  }
}

void launch_reduce_q(int nparts, int Q, const QPart* part, QStat* out, hipStream_t st, const double* cnull) {
  if (Q == 0) return;
  hipLaunchKernelGGL(k_reduce_q, dim3(Q), dim3(256), 0, st, nparts, Q, part, cnull, out);
}

// Isolated-vertex columns: one pass over X[n][P] giving per-block partials of LSE(-X) and of (min X, lowest row).
__global__ __launch_bounds__(256) void k_iso_cols(int64_t n, int P, const float* __restrict__ X, Lse* __restrict__ lpart,
                                                  float* __restrict__ pv, int* __restrict__ pi) {
  __shared__ Lse red[256];
  __shared__ float rv[256];
  __shared__ int ri[256];
  const int t = threadIdx.x, RP = 256 / P;
  const bool active = t < RP * P;
  const int p = active ? t % P : 0;
  Lse acc = lse_empty();
  float best = INFINITY;
  int bi = INT_MAX;
  if (active)
    for (int64_t v = (int64_t)blockIdx.x * RP + t / P; v < n; v += (int64_t)gridDim.x * RP) {
      const float x = X[v * P + p];
      lse_add(acc, -(double)x);
      if (x < best) {  // rows ascend within a thread
        best = x;
        bi = (int)v;
      }
    }
  red[t] = acc;
  rv[t] = best;
  ri[t] = bi;
  __syncthreads();
  if (t < P) {
    Lse s = lse_empty();
    for (int r = 0; r < RP; ++r) {
      s = lse_merge(s, red[r * P + t]);
      const float x = rv[r * P + t];
      const int xi = ri[r * P + t];
      if (x < best || (x == best && xi < bi)) {
        best = x;
        bi = xi;
      }
    }
    lpart[(int64_t)blockIdx.x * P + t] = s;
    pv[(int64_t)blockIdx.x * P + t] = best;
    pi[(int64_t)blockIdx.x * P + t] = bi;
  }
}

// One block per column: partials merged by a fixed tree (the minimum's lowest row wins ties in any order).
__global__ __launch_bounds__(256) void k_iso_final(int nparts, int P, int base, const Lse* __restrict__ lpart,
                                                   const float* __restrict__ pv, const int* __restrict__ pi,
                                                   IsoCol* __restrict__ out) {
  __shared__ Lse rs[256];
  __shared__ float rv[256];
  __shared__ int ri[256];
  const int p = blockIdx.x, t = threadIdx.x;
  Lse s = lse_empty();
  float best = INFINITY;
  int bi = INT_MAX;
  for (int b = t; b < nparts; b += 256) {
    s = lse_merge(s, lpart[(int64_t)b * P + p]);
    const float x = pv[(int64_t)b * P + p];
    const int xi = pi[(int64_t)b * P + p];
    if (x < best || (x == best && xi < bi)) {
      best = x;
      bi = xi;
    }
  }
  rs[t] = s;
  rv[t] = best;
  ri[t] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      rs[t] = lse_merge(rs[t], rs[t + w]);
      if (rv[t + w] < rv[t] || (rv[t + w] == rv[t] && ri[t + w] < ri[t])) {
        rv[t] = rv[t + w];
        ri[t] = ri[t + w];
      }
    }
    __syncthreads();
  }
  // A column whose isolated costs are all +inf has no finite share: treat it as empty.
  if (t == 0)
    out[p] = ri[0] == INT_MAX ? IsoCol{-kInf, 0.0, INFINITY, -1} : IsoCol{rs[0].m, rs[0].t, rv[0], base + ri[0]};
}

void launch_iso_cols(int64_t n, int P, const float* X, int base, Lse* lpart, float* pv, int* pi, int nparts,
                     IsoCol* out, hipStream_t st) {
  if (P == 0) return;
  hipLaunchKernelGGL(k_iso_cols, dim3(nparts), dim3(256), 0, st, n, P, X, lpart, pv, pi);
  hipLaunchKernelGGL(k_iso_final, dim3(P), dim3(256), 0, st, nparts, P, base, lpart, pv, pi, out);
}

// Scalars of the prep state (src/main.cu:116-146): MU0 = LSM(-CV), RE0 = FE0 = LSM(-CE).  The column maximum of
// MU0 sits at the minimum cost, where (x - m) = 0, so max_v MU0[v][p] = -log1p(t_p).
__global__ void k_prep_scalars(int Q, const int* __restrict__ ps, const int* __restrict__ pd,
                               const Lse2* __restrict__ cv, const Lse2* __restrict__ ce, QStat* __restrict__ qs) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  qs[q] = QStat{-cv[ps[q]].l, -cv[pd[q]].l, ce[q].m, ce[q].l, ce[q].m, ce[q].l, -1, -1, kInf};
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

// -inf into three ranges of each of nslots slots (stride `stride` floats): [a0, a0 + al) of A, and [b0, b0 + bl) and
// [c0, c0 + cl) of B.  One launch for every slot.
__global__ void k_fill_slots(float* __restrict__ A, float* __restrict__ B, int64_t nslots, int64_t stride, int64_t a0,
                             int64_t al, int64_t b0, int64_t bl, int64_t c0, int64_t cl) {
  const int64_t per = al + bl + cl, total = nslots * per;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i / per, r = i - s * per;
    if (r < al)
      A[s * stride + a0 + r] = -INFINITY;
    else if (r < al + bl)
      B[s * stride + b0 + (r - al)] = -INFINITY;
    else
      B[s * stride + c0 + (r - al - bl)] = -INFINITY;
  }
}

void launch_fill_slots(float* A, float* B, int64_t nslots, int64_t stride, int64_t a0, int64_t al, int64_t b0,
                       int64_t bl, int64_t c0, int64_t cl, hipStream_t st) {
  const int64_t total = nslots * (al + bl + cl);
  if (total == 0) return;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(k_fill_slots, dim3(blocks), dim3(256), 0, st, A, B, nslots, stride, a0, al, b0, bl, c0, cl);
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
    Lse s[4] = {lse_empty(), lse_empty(), lse_empty(), lse_empty()};
    int cur = ks[R.kb].y;
    // A finished source vertex.  It runs at vertex changes, which differ between the wave's runs, so each costs the
    // whole wave: its log1p and the merge into the block partial take FP32 transcendentals (lsO is stored in FP32;
    // the merge's exponential has the precision of lse_add's terms).
    auto emit = [&](int v) {
      float4 mn, ls, vf;
      mn.x = (float)(-s[0].m), mn.y = (float)(-s[1].m), mn.z = (float)(-s[2].m), mn.w = (float)(-s[3].m);
      ls.x = log1pf((float)s[0].t), ls.y = log1pf((float)s[1].t), ls.z = log1pf((float)s[2].t);
      ls.w = log1pf((float)s[3].t);
      vf.x = (float)s[0].m, vf.y = (float)s[1].m, vf.z = (float)s[2].m, vf.w = (float)s[3].m;
      reinterpret_cast<float4*>(minCEo)[(int64_t)v * L4 + j] = mn;
      reinterpret_cast<float4*>(lsO)[(int64_t)v * L4 + j] = ls;
      reinterpret_cast<float4*>(VFm)[(int64_t)v * L4 + j] = vf;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        glob[c] = lse_merge_fast(glob[c], s[c]);
        s[c] = lse_empty();
      }
    };
    // Steps in batches of kSrcBatch: the batch's (row, vertex) pairs, then its cost rows, are loaded together (two
    // memory round trips per batch rather than per step); the steps are then taken in order.  Slots past the run's
    // end load its last pair (a valid address) and are not taken.
    constexpr int kSrcBatch = 8;
    for (int k0 = R.kb; k0 < R.ke; k0 += kSrcBatch) {
      int2 e[kSrcBatch];
      float4 c[kSrcBatch];
#pragma unroll
      for (int u = 0; u < kSrcBatch; ++u) e[u] = ks[k0 + u < R.ke ? k0 + u : R.ke - 1];
#pragma unroll
      for (int u = 0; u < kSrcBatch; ++u) c[u] = reinterpret_cast<const float4*>(CE + ce_index(e[u].x, 4 * j, Qs))[0];
#pragma unroll
      for (int u = 0; u < kSrcBatch; ++u) {
        if (k0 + u >= R.ke) break;
        if (e[u].y != cur) {
          emit(cur);
          cur = e[u].y;
        }
        lse_add(s[0], -(double)c[u].x);
        lse_add(s[1], -(double)c[u].y);
        lse_add(s[2], -(double)c[u].z);
        lse_add(s[3], -(double)c[u].w);
      }
    }
    if (R.slot >= 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        carry[((int64_t)R.slot * Qs + 4 * j + c) * 2 + 0] = s[c].m;
        carry[((int64_t)R.slot * Qs + 4 * j + c) * 2 + 1] = s[c].t;
        glob[c] = lse_merge_fast(glob[c], s[c]);
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

// One wave per split vertex: groups of Qs/4 lanes take its pieces round-robin (lane j owns pattern edges
// 4j..4j+3), then the groups' results are merged in group order through LDS.  Deterministic.
__global__ __launch_bounds__(256) void k_walk_src_fixup(int nsplit, const Split* __restrict__ splits,
                                                        const double* __restrict__ carry, int Q, int Qs,
                                                        float* __restrict__ minCEo, float* __restrict__ lsO,
                                                        float* __restrict__ VFm) {
  __shared__ Lse red[4][64][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L4 = Qs / 4, G = 64 / L4;
  const int g = lane / L4, j = lane - g * L4;
  const int sidx = blockIdx.x * 4 + wave;
  const bool active = sidx < nsplit && g < G;
  Lse s[4] = {lse_empty(), lse_empty(), lse_empty(), lse_empty()};
  Split S{0, 0, 0, 0};
  if (sidx < nsplit) S = splits[sidx];
  if (active)
    for (int c = g; c < S.count; c += G) {
      const double* o = carry + ((int64_t)(S.first + c) * Qs + 4 * j) * 2;
#pragma unroll
      for (int u = 0; u < 4; ++u) s[u] = lse_merge(s[u], Lse{o[2 * u], o[2 * u + 1]});
    }
#pragma unroll
  for (int u = 0; u < 4; ++u) red[wave][lane][u] = s[u];
  __syncthreads();
  if (sidx < nsplit && g == 0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      Lse acc = lse_empty();
      for (int gg = 0; gg < G; ++gg) acc = lse_merge(acc, red[wave][gg * L4 + j][u]);
      const int q = 4 * j + u;
      const int64_t o = (int64_t)S.v * Qs + q;
      minCEo[o] = q < Q ? (float)(-acc.m) : 0.f;
      lsO[o] = q < Q ? (float)log1p(acc.t) : 0.f;
      VFm[o] = q < Q ? (float)acc.m : 0.f;
    }
  }
}

void launch_walk_src_fixup(int nsplit, const Split* splits, const double* carry, int Q, int Qs, float* minCEo,
                           float* lsO, float* VFm, hipStream_t st) {
  if (nsplit == 0) return;
  hipLaunchKernelGGL(k_walk_src_fixup, dim3((unsigned)((nsplit + 3) / 4)), dim3(256), 0, st, nsplit, splits, carry,
                     Q, Qs, minCEo, lsO, VFm);
}

// ---------------------------------------------------------------------------------------------------------------
// Destination maxima (CSC walk), src/main.cu:220-224 with the log-softmax of src/main.cu:218 folded in:
//   FMaxN[v][q] = max over in-edges e of v of (VR[src e][q] - CE[e][q])  -  log sum_e exp(FE[q][e])
// computed exactly in FP64 (a difference of two floats is exact in double) and rounded once.  With VR == nullptr
// it is the prep's max over in-edges of -CE (src/main.cu:143-146).
//
// One wave walks one bundle of G runs, group g (Qs/4 lanes) walking run g, lane j owning pattern edges 4j..4j+3.
// Step s of the bundle reads rows base + s*G + g of CE (one contiguous block of G*Qs floats for the wave), of bsrc
// and of bdst, and gathers row bsrc of VR.  The steps go in batches of B, software-pipelined two batches deep with
// two register sets used in turn (no register copies, so no wait for loads in flight at a loop back-edge): while
// batch t is reduced, the costs, destinations and gathered rows of batch t + 1 and the sources of batches t + 2 and
// t + 3 are in flight.  Loads past the bundle's last batch are clamped to it and never used.

int walk_groups(int Qs) { return 64 / (Qs / 4); }

int walk_batch() {
  if (const char* e = std::getenv("AC_WALK_BATCH")) {  // tuning override: 1, 2 or 4
    const int b = std::atoi(e);
    if (b == 1 || b == 2 || b == 4) return b;
  }
  return 4;  // measured at R-MAT 22 with packed costs and runs of 128 edges
}

// kArg: also records, per (v, q), the row of the in-edge attaining the maximum (the first one in walk order on
// ties; -1 if none) into FA[v][Qs], and per carry slot into carg, for the adjoint's max-subgradient routing.
// kPrune (bmin != nullptr): exact pruning, per bundle and per lane (below).
template <bool kWithVR, int B, bool kArg, bool kPrune>
__global__ __launch_bounds__(kWalkBlock) void k_walk_max(int64_t nbundles, const Bundle* __restrict__ bundles,
                                                         const int* __restrict__ gslot, const int* __restrict__ bsrc,
                                                         const int* __restrict__ bdst, const float* __restrict__ CE,
                                                         const float* __restrict__ VR, int VRs, int Q, int Qs,
                                                         const QStat* __restrict__ qs, float* __restrict__ FMaxN,
                                                         double* __restrict__ carry, int* __restrict__ FA,
                                                         int* __restrict__ carg, const float* __restrict__ bmin,
                                                         const unsigned* __restrict__ CEp,
                                                         const unsigned* __restrict__ bpack,
                                                         unsigned* __restrict__ CEpo) {
  const int L4 = Qs / 4, G = 64 / L4, VR4s = VRs / 4;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane / L4, j = lane - g * L4;
  const int64_t b = (int64_t)blockIdx.x * (kWalkBlock / 64) + wave;  // wave-uniform: the bundle header is scalar
  if (g >= G || b >= nbundles) return;
  const Bundle Bd = bundles[b];
  const int slot = gslot[b * G + g];
  // Exact pruning: an edge whose cost is at or above its pattern edge's bound cannot raise FMax above the fill, so a
  // term that is -inf instead gives the same FMax (the fill wins, as it would have).  A wave with no live lane skips
  // its bundle and writes -inf rows; a dead lane (none of its four pattern edges has a cost below the bound in this
  // bundle) skips its cost loads and VR gathers, and its maxima stay -inf.
  bool live = true;
  if (kPrune) {
    live = false;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = 4 * j + u;
      if (q < Q && (double)bmin[b * Qs + q] < qs[q].cthr) live = true;
    }
    if (!__any(live)) {
      if (slot >= 0 && kArg) {
        double* o = carry + (int64_t)slot * Qs + 4 * j;
        o[0] = o[1] = o[2] = o[3] = -kInf;
      } else if (slot >= 0) {
        reinterpret_cast<f4v*>(carry)[(int64_t)slot * L4 + j] = f4v{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
      } else {
        const int v0 = bdst[Bd.base + g], v1 = bdst[Bd.base + (int64_t)(Bd.len - 1) * G + g];
        const float4 ninf = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
        if (v0 >= 0)
          for (int v = v0; v <= v1; ++v) reinterpret_cast<float4*>(FMaxN)[(int64_t)v * L4 + j] = ninf;
      }
      return;
    }
  }
  // log sum_e exp(FE[q][e]) as one FP64 value (its FP64 rounding is far below the FP32 rounding of the result).
  double sl[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int q = 4 * j + c;
    sl[c] = q < Q ? qs[q].mFE + qs[q].lFE : 0.0;
  }
  double mx0 = -kInf, mx1 = -kInf, mx2 = -kInf, mx3 = -kInf;
  int a0 = -1, a1 = -1, a2 = -1, a3 = -1;
  int cur = -1;
  // Finished vertex v: its row of maxima, normalised (a macro rather than a lambda, so that the walk's loop-carried
  // state is not captured by reference twice over and stays in registers).
#define AC_WK_EMIT(V)                                                                           \
  {                                                                                             \
    const f4v o_ = f4v{(float)(mx0 - sl[0]), (float)(mx1 - sl[1]), (float)(mx2 - sl[2]),         \
                       (float)(mx3 - sl[3])};                                                   \
    reinterpret_cast<f4v*>(FMaxN)[(int64_t)(V) * L4 + j] = o_;                                  \
    if (kArg) {                                                                                 \
      reinterpret_cast<int4*>(FA)[(int64_t)(V) * L4 + j] = make_int4(a0, a1, a2, a3);           \
      a0 = a1 = a2 = a3 = -1;                                                                   \
    }                                                                                           \
    mx0 = mx1 = mx2 = mx3 = -kInf;                                                              \
  }
  const int GL4 = G * L4;
  // Streams: a wave-uniform bundle base plus small per-lane offsets (32-bit address arithmetic).
  const int* srcb = bsrc + Bd.base;
  const f4v* ceb = reinterpret_cast<const f4v*>(CE) + Bd.base * L4;
  const unsigned so = (unsigned)g, co = (unsigned)(g * L4 + j);
  // Plain walk (kWithVR, !kArg): destinations come from the sources' new-vertex flags (a run's vertices are
  // consecutive), so only the run's first destination is read.
  constexpr bool kFlags = kWithVR && !kArg;
  if (kFlags) cur = bdst[Bd.base + g];
  const int* dstp = bdst + Bd.base + g;
  const f4v* VRv = reinterpret_cast<const f4v*>(VR);
  const int last = Bd.len / B - 1;
  // Packed costs (CEp != nullptr): the bundle's base, 0 when it keeps its FP32 rows.  Decoding is exact, so maxima
  // and recorded argmax rows are those of the FP32 walk.
  const unsigned pbase = CEp ? bpack[b] : 0u;
  const unsigned* cpb = CEp ? CEp + Bd.base * L4 * 3 : nullptr;
  // Packing on the way (the prep walk, CEpo != nullptr, no pruning): every row is reduced exactly once from its FP32
  // costs, and is also written out as 24-bit codes when the bundle's base is nonzero (pack_base: each cost of a column
  // q < Q other than +inf codes as its bit pattern minus the base, +inf and columns q >= Q as 0).
  const unsigned obase = CEpo && !kPrune ? bpack[b] : 0u;
  unsigned* const cpo = obase ? CEpo + Bd.base * L4 * 3 : nullptr;
#define AC_WK_SRC(S, T)                                                          \
  if (kWithVR) {                                                                 \
    const int t_ = (T) < last ? (T) : last;                                      \
    _Pragma("unroll") for (int u = 0; u < B; ++u) S[u] =                         \
        __builtin_nontemporal_load(srcb + (so + (unsigned)((t_ * B + u) * G)));  \
  }
#define AC_WK_BATCH(C, D, W, S, T)                                                                       \
  {                                                                                                      \
    const int t_ = (T) < last ? (T) : last;                                                              \
    _Pragma("unroll") for (int u = 0; u < B; ++u) {                                                      \
      if constexpr (kP)                                                                                  \
        C[u] = (!kPrune || live) ? nt_load_u3(cpb + 3 * (co + (unsigned)((t_ * B + u) * GL4)))           \
                                 : U3{0u, 0u, 0u};                                                       \
      else                                                                                               \
        C[u] = (!kPrune || live) ? __builtin_nontemporal_load(ceb + (co + (unsigned)((t_ * B + u) * GL4))) \
                                 : f4v{INFINITY, INFINITY, INFINITY, INFINITY};                          \
    }                                                                                                    \
    _Pragma("unroll") for (int u = 0; u < B; ++u) D[u] =                                                 \
        kFlags ? (S[u] & kNewVertex) : __builtin_nontemporal_load(dstp + (t_ * B + u) * G);              \
    _Pragma("unroll") for (int u = 0; u < B; ++u) W[u] =                                                 \
        (kWithVR && (!kPrune || live)) ? VRv[(int64_t)(S[u] & ~kNewVertex) * VR4s + j]                   \
                                       : f4v{0.f, 0.f, 0.f, 0.f};                                        \
  }
#define AC_WK_REDUCE(C, D, W, T)                                                    \
  _Pragma("unroll") for (int u = 0; u < B; ++u) {                                   \
    if (kFlags) {                                                                   \
      if (D[u]) {                                                                   \
        AC_WK_EMIT(cur);                                                            \
        ++cur;                                                                      \
      }                                                                             \
    } else if (D[u] != cur) {                                                       \
      if (cur >= 0) AC_WK_EMIT(cur);                                                \
      cur = D[u];                                                                   \
    }                                                                               \
    f4v c_;                                                                         \
    if constexpr (kP)                                                               \
      c_ = decode_costs(C[u], pbase);                                               \
    else                                                                            \
      c_ = C[u];                                                                    \
    if (!kP && cpo) {                                                               \
      unsigned k_[4];                                                               \
      _Pragma("unroll") for (int w_ = 0; w_ < 4; ++w_) {                            \
        const unsigned x_ = __float_as_uint(c_[w_]);                                \
        k_[w_] = (4 * j + w_ < Q && x_ != 0x7F800000u) ? x_ - obase : 0u;           \
      }                                                                             \
      unsigned* o_ = cpo + 3 * (co + (unsigned)(((T) * B + u) * GL4));             \
      o_[0] = k_[0] | (k_[1] << 24);                                                \
      o_[1] = (k_[1] >> 8) | (k_[2] << 16);                                         \
      o_[2] = (k_[2] >> 16) | (k_[3] << 8);                                         \
    }                                                                               \
    const double d0 = (double)W[u].x - (double)c_.x;                                \
    const double d1 = (double)W[u].y - (double)c_.y;                                \
    const double d2 = (double)W[u].z - (double)c_.z;                                \
    const double d3 = (double)W[u].w - (double)c_.w;                                \
    if (kArg) {                                                                     \
      const int row_ = (int)(Bd.base + g) + ((T) * B + u) * G;                       \
      if (d0 > mx0) mx0 = d0, a0 = row_;                                            \
      if (d1 > mx1) mx1 = d1, a1 = row_;                                            \
      if (d2 > mx2) mx2 = d2, a2 = row_;                                            \
      if (d3 > mx3) mx3 = d3, a3 = row_;                                            \
    } else {                                                                        \
      mx0 = fmax(mx0, d0);                                                          \
      mx1 = fmax(mx1, d1);                                                          \
      mx2 = fmax(mx2, d2);                                                          \
      mx3 = fmax(mx3, d3);                                                          \
    }                                                                               \
  }
  // The pipelined walk over the bundle, for FP32 cost rows (kP false) or packed ones (decoded when reduced, so the
  // loads stay in flight as raw dwords).  kP is written in terms of the template parameters so that it is
  // value-dependent: the discarded `if constexpr` branches are then not instantiated.
#define AC_WK_PIPELINE(KP)                                                   \
  {                                                                          \
    constexpr bool kP = (KP);                                                \
    typedef typename std::conditional<kP, U3, f4v>::type CostRow;            \
    int SA[B], SB[B], DA[B], DB[B];                                          \
    CostRow CA[B], CB[B];                                                    \
    f4v WA[B], WB[B];                                                        \
    _Pragma("unroll") for (int u = 0; u < B; ++u) SA[u] = SB[u] = 0;         \
    AC_WK_SRC(SA, 0);                                                        \
    AC_WK_SRC(SB, 1);                                                        \
    AC_WK_BATCH(CA, DA, WA, SA, 0);                                          \
    AC_WK_SRC(SA, 2);                                                        \
    for (int t = 0;; t += 2) {                                               \
      AC_WK_BATCH(CB, DB, WB, SB, t + 1);                                    \
      AC_WK_SRC(SB, t + 3);                                                  \
      AC_WK_REDUCE(CA, DA, WA, t);                                           \
      if (t + 1 > last) break;                                               \
      AC_WK_BATCH(CA, DA, WA, SA, t + 2);                                    \
      AC_WK_SRC(SA, t + 4);                                                  \
      AC_WK_REDUCE(CB, DB, WB, t + 1);                                       \
      if (t + 2 > last) break;                                               \
    }                                                                        \
  }
  if (pbase)
    AC_WK_PIPELINE(kWithVR || B > 0)
  else
    AC_WK_PIPELINE(false && kWithVR)
#undef AC_WK_PIPELINE
#undef AC_WK_SRC
#undef AC_WK_BATCH
#undef AC_WK_REDUCE
  if (slot >= 0 && kArg) {
    double* o = carry + (int64_t)slot * Qs + 4 * j;
    o[0] = mx0;
    o[1] = mx1;
    o[2] = mx2;
    o[3] = mx3;
    reinterpret_cast<int4*>(carg)[(int64_t)slot * L4 + j] = make_int4(a0, a1, a2, a3);
  } else if (slot >= 0) {
    // A piece of a split vertex: its maximum, normalised and rounded like a finished row.  Rounding is monotone,
    // so the maximum over the pieces of the rounded values is the rounded maximum (exact).
    reinterpret_cast<f4v*>(carry)[(int64_t)slot * L4 + j] = f4v{(float)(mx0 - sl[0]), (float)(mx1 - sl[1]),
                                                                (float)(mx2 - sl[2]), (float)(mx3 - sl[3])};
  } else if (cur >= 0) {
    AC_WK_EMIT(cur);
  }
#undef AC_WK_EMIT
}

template <bool kWithVR, bool kArg, bool kPrune>
static void launch_walk_max_t(int B, int nb, hipStream_t st, int64_t nbundles, const Bundle* bundles,
                              const int* gslot, const int* bsrc, const int* bdst, const float* CE, const float* VR,
                              int VRs, int Q, int Qs, const QStat* qs, float* FMaxN, double* carry, int* FA,
                              int* carg, const float* bmin, const unsigned* CEp, const unsigned* bpack,
                              unsigned* CEpo) {
  switch (B) {
    case 1:
      hipLaunchKernelGGL((k_walk_max<kWithVR, 1, kArg, kPrune>), dim3(nb), dim3(kWalkBlock), 0, st, nbundles,
                         bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs, FMaxN, carry, FA, carg, bmin, CEp, bpack,
                         CEpo);
      break;
    case 4:
      hipLaunchKernelGGL((k_walk_max<kWithVR, 4, kArg, kPrune>), dim3(nb), dim3(kWalkBlock), 0, st, nbundles,
                         bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs, FMaxN, carry, FA, carg, bmin, CEp, bpack,
                         CEpo);
      break;
    default:
      hipLaunchKernelGGL((k_walk_max<kWithVR, 2, kArg, kPrune>), dim3(nb), dim3(kWalkBlock), 0, st, nbundles,
                         bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs, FMaxN, carry, FA, carg, bmin, CEp, bpack,
                         CEpo);
      break;
  }
}

template <bool kWithVR, bool kArg>
static void launch_walk_max_p(int B, int nb, hipStream_t st, int64_t nbundles, const Bundle* bundles,
                              const int* gslot, const int* bsrc, const int* bdst, const float* CE, const float* VR,
                              int VRs, int Q, int Qs, const QStat* qs, float* FMaxN, double* carry, int* FA,
                              int* carg, const float* bmin, const unsigned* CEp, const unsigned* bpack,
                              unsigned* CEpo) {
  if (bmin)
    launch_walk_max_t<kWithVR, kArg, true>(B, nb, st, nbundles, bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs,
                                           FMaxN, carry, FA, carg, bmin, CEp, bpack, CEpo);
  else
    launch_walk_max_t<kWithVR, kArg, false>(B, nb, st, nbundles, bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs,
                                            FMaxN, carry, FA, carg, bmin, CEp, bpack, CEpo);
}

void launch_walk_max(int64_t nbundles, const Bundle* bundles, const int* gslot, const int* bsrc, const int* bdst,
                     const float* CE, const float* VR, int VRs, int Q, int Qs, const QStat* qs, float* FMaxN,
                     double* carry, int B, hipStream_t st, int* FA, int* carg, const float* bmin,
                     const unsigned* CEp, const unsigned* bpack, unsigned* CEp_out) {
  if (nbundles == 0) return;
  const int nb = (int)((nbundles + kWalkBlock / 64 - 1) / (kWalkBlock / 64));
  if (!bpack) CEp = CEp_out = nullptr;
  if (CEp_out) CEp = nullptr;  // packing on the way: the walk reads the FP32 rows
  if (VR && FA)
    launch_walk_max_p<true, true>(B, nb, st, nbundles, bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs, FMaxN,
                                  carry, FA, carg, bmin, CEp, bpack, CEp_out);
  else if (VR)
    launch_walk_max_p<true, false>(B, nb, st, nbundles, bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs, FMaxN,
                                   carry, FA, carg, bmin, CEp, bpack, CEp_out);
  else if (FA)
    launch_walk_max_p<false, true>(B, nb, st, nbundles, bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs, FMaxN,
                                   carry, FA, carg, bmin, CEp, bpack, CEp_out);
  else
    launch_walk_max_p<false, false>(B, nb, st, nbundles, bundles, gslot, bsrc, bdst, CE, VR, VRs, Q, Qs, qs, FMaxN,
                                    carry, FA, carg, bmin, CEp, bpack, CEp_out);
}

// The base of a bundle's 24-bit cost codes from its least and largest cost bit patterns over columns q < Q, +inf
// excluded (lo > hi: no such cost).  The costs must lie strictly inside two binades (2^e, 2^(e+2)) of one biased
// exponent e >= 1: then each is its bit pattern minus base = e << 23, a code in [1, 2^24), and +inf (padding rows)
// codes as 0.  Positive floats order as their bit patterns, so the least and largest patterns decide: a least cost
// that is a power of two widens the window down one binade, and negative, zero, subnormal or NaN costs fall outside
// every window.  0 = the bundle keeps its FP32 rows.
__device__ __forceinline__ unsigned pack_base(unsigned lo, unsigned hi) {
  if (lo > hi || hi >= 0x7F800000u) return 0u;
  unsigned e = lo >> 23;
  if ((lo & 0x7FFFFFu) == 0u) e = e > 0u ? e - 1u : 0u;  // code 0 is +inf: the least cost must lie above the base
  return e >= 1u && hi - (e << 23) < (1u << 24) ? e << 23 : 0u;
}

// Per bundle and pattern edge, the least cost among the bundle's rows (padding rows are +inf): bmin[b][Qs].  One
// wave per bundle, as the walk; the groups' minima are merged through LDS.  With bpack != nullptr it also writes
// the bundle's cost-code base (pack_base), from the same pass over the rows.
__global__ __launch_bounds__(kWalkBlock) void k_bundle_min(int64_t nbundles, const Bundle* __restrict__ bundles,
                                                           const float* __restrict__ CE, int Q, int Qs,
                                                           float* __restrict__ bmin, unsigned* __restrict__ bpack) {
  __shared__ f4v red[kWalkBlock / 64][64];
  const int L4 = Qs / 4, G = 64 / L4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane / L4, j = lane - g * L4;
  const int64_t b = (int64_t)blockIdx.x * (kWalkBlock / 64) + wave;
  f4v m = f4v{INFINITY, INFINITY, INFINITY, INFINITY};
  unsigned lo = 0xFFFFFFFFu, hi = 0u;
  if (g < G && b < nbundles) {
    const Bundle Bd = bundles[b];
    const f4v* ce = reinterpret_cast<const f4v*>(CE) + (Bd.base + g) * L4 + j;
    for (int s = 0; s < Bd.len; ++s) {
      const f4v c = __builtin_nontemporal_load(ce + (int64_t)s * G * L4);
      m = f4v{fminf(m[0], c[0]), fminf(m[1], c[1]), fminf(m[2], c[2]), fminf(m[3], c[3])};
      if (bpack) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const unsigned x = __float_as_uint(c[u]);
          if (4 * j + u < Q && x != 0x7F800000u) {
            lo = min(lo, x);
            hi = max(hi, x);
          }
        }
      }
    }
  }
  if (bpack) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo = min(lo, (unsigned)__shfl_xor((int)lo, o, 64));
      hi = max(hi, (unsigned)__shfl_xor((int)hi, o, 64));
    }
    if (lane == 0 && b < nbundles) bpack[b] = pack_base(lo, hi);
  }
  red[wave][lane] = m;
  __syncthreads();
  if (b < nbundles && lane < L4) {
    f4v a = red[wave][lane];
    for (int gg = 1; gg < G; ++gg) {
      const f4v c = red[wave][gg * L4 + lane];
      a = f4v{fminf(a[0], c[0]), fminf(a[1], c[1]), fminf(a[2], c[2]), fminf(a[3], c[3])};
    }
    reinterpret_cast<f4v*>(bmin)[b * L4 + lane] = a;
  }
}

void launch_bundle_min(int64_t nbundles, const Bundle* bundles, const float* CE, int Q, int Qs, float* bmin,
                       unsigned* bpack, hipStream_t st) {
  if (nbundles == 0) return;
  const int nb = (int)((nbundles + kWalkBlock / 64 - 1) / (kWalkBlock / 64));
  hipLaunchKernelGGL(k_bundle_min, dim3(nb), dim3(kWalkBlock), 0, st, nbundles, bundles, CE, Q, Qs, bmin, bpack);
}

// Bundle layout from the CSC arrays: one block per bundle writes its rows (step-major, run-minor).  rowmap[k] is
// the row of CSC position k.  brun[b * G + g] is the run walked by group g of bundle b (-1: none).
__global__ __launch_bounds__(256) void k_build_bundles(const Bundle* __restrict__ bundles, const int* __restrict__ brun,
                                                       const Run* __restrict__ runs, const int2* __restrict__ sd,
                                                       const int* __restrict__ eid, int G, int* __restrict__ bsrc,
                                                       int* __restrict__ bdst, int* __restrict__ beid,
                                                       int* __restrict__ rowmap) {
  const int64_t b = blockIdx.x;
  const Bundle Bd = bundles[b];
  const int nrow = Bd.len * G;
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
    const int s = i / G, g = i - s * G;
    const int r = brun[b * G + g];
    const int64_t row = Bd.base + i;
    int src = 0, dst = -1, e = -1;
    if (r >= 0) {
      const Run R = runs[r];
      if (R.kb + s < R.ke) {
        const int k = R.kb + s;
        src = sd[k].x;
        dst = sd[k].y;
        e = eid[k];
        rowmap[k] = (int)row;
        if (s > 0 && sd[k - 1].y != dst) src |= kNewVertex;  // the run moves on to the next vertex (dst + 1)
      } else {
        dst = sd[R.ke - 1].y;
      }
    }
    bsrc[row] = src;
    bdst[row] = dst;
    beid[row] = e;
  }
}

void launch_build_bundles(int64_t nbundles, const Bundle* bundles, const int* brun, const Run* runs, const int2* sd,
                          const int* eid, int G, int* bsrc, int* bdst, int* beid, int* rowmap, hipStream_t st) {
  if (nbundles == 0) return;
  hipLaunchKernelGGL(k_build_bundles, dim3((unsigned)nbundles), dim3(256), 0, st, bundles, brun, runs, sd, eid, G,
                     bsrc, bdst, beid, rowmap);
}

__global__ void k_remap_ks(int64_t m, const int2* __restrict__ ks, const int* __restrict__ rowmap,
                           int2* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int2 e = ks[i];
    out[i] = make_int2(rowmap[e.x], e.y);
  }
}

void launch_remap_ks(int64_t m, const int2* ks, const int* rowmap, int2* out, hipStream_t st) {
  if (m == 0) return;
  const int blocks = (int)std::min<int64_t>((m + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(k_remap_ks, dim3(blocks), dim3(256), 0, st, m, ks, rowmap, out);
}

// One wave per split vertex, as k_walk_src_fixup.  With carg: the argmax row of the best piece (the first in merge
// order on ties) into FA.
__global__ __launch_bounds__(256) void k_walk_max_fixup(int nsplit, const Split* __restrict__ splits,
                                                        const double* __restrict__ carry, int Q, int Qs,
                                                        const QStat* __restrict__ qs, float* __restrict__ FMaxN,
                                                        const int* __restrict__ carg, int* __restrict__ FA) {
  __shared__ double red[4][64][4];
  __shared__ int rarg[4][64][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L4 = Qs / 4, G = 64 / L4;
  const int g = lane / L4, j = lane - g * L4;
  const int sidx = blockIdx.x * 4 + wave;
  const bool active = sidx < nsplit && g < G;
  double mx[4] = {-kInf, -kInf, -kInf, -kInf};
  int ag[4] = {-1, -1, -1, -1};
  Split S{0, 0, 0, 0};
  if (sidx < nsplit) S = splits[sidx];
  if (active)
    for (int c = g; c < S.count; c += G) {
      const double4 o = *reinterpret_cast<const double4*>(carry + (int64_t)(S.first + c) * Qs + 4 * j);
      const double ov[4] = {o.x, o.y, o.z, o.w};
      int4 oa = make_int4(-1, -1, -1, -1);
      if (carg) oa = reinterpret_cast<const int4*>(carg)[(int64_t)(S.first + c) * L4 + j];
      const int oav[4] = {oa.x, oa.y, oa.z, oa.w};
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (ov[u] > mx[u]) {
          mx[u] = ov[u];
          ag[u] = oav[u];
        }
    }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    red[wave][lane][u] = mx[u];
    rarg[wave][lane][u] = ag[u];
  }
  __syncthreads();
  if (sidx < nsplit && g == 0) {
    float4 out;
    float* po = reinterpret_cast<float*>(&out);
    int oarg[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double acc = -kInf;
      int aa = -1;
      for (int gg = 0; gg < G; ++gg)
        if (red[wave][gg * L4 + j][u] > acc) {
          acc = red[wave][gg * L4 + j][u];
          aa = rarg[wave][gg * L4 + j][u];
        }
      const int q = 4 * j + u;
      // Normalised as the walk normalises a vertex it finishes (one FP64 subtraction of mFE + lFE, rounded once),
      // so a recording walk leaves the same FMaxN as a plain one, whose pieces are normalised that way.
      po[u] = (float)(acc - (q < Q ? qs[q].mFE + qs[q].lFE : 0.0));
      oarg[u] = aa;
    }
    reinterpret_cast<float4*>(FMaxN)[(int64_t)S.v * L4 + j] = out;
    if (carg) reinterpret_cast<int4*>(FA)[(int64_t)S.v * L4 + j] = make_int4(oarg[0], oarg[1], oarg[2], oarg[3]);
  }
}

// Plain path: the pieces hold normalised FP32 maxima, and their maximum is exact in any order.  Pieces are grouped
// into chunks of at most kFixChunk (Split{v, first piece, count, out}); one group of Qs/4 lanes takes a chunk with
// all its loads in flight and writes the vertex's row (out < 0) or chunk partial `out`; vertices of several chunks
// (MSplit{v, first partial, count}) are then finished from their partials, one wave each.
__global__ __launch_bounds__(256) void k_fix_chunks(int nchunk, const Split* __restrict__ chunks,
                                                    const f4v* __restrict__ carry, int Qs, f4v* __restrict__ part,
                                                    f4v* __restrict__ FMaxN) {
  const int lane = threadIdx.x & 63;
  const int L4 = Qs / 4, G = 64 / L4;
  const int g = lane / L4, j = lane - g * L4;
  if (g >= G) return;
  const int64_t ngroups = (int64_t)gridDim.x * (blockDim.x / 64) * G;
  for (int64_t ci = ((int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * G + g; ci < nchunk;
       ci += ngroups) {
    const Split C = chunks[ci];
    const f4v* c = carry + (int64_t)C.first * L4 + j;
    f4v x[kFixChunk];
#pragma unroll
    for (int k = 0; k < kFixChunk; ++k)
      x[k] = k < C.count ? c[(int64_t)k * L4] : f4v{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    f4v m = x[0];
#pragma unroll
    for (int k = 1; k < kFixChunk; ++k)
      m = f4v{fmaxf(m.x, x[k].x), fmaxf(m.y, x[k].y), fmaxf(m.z, x[k].z), fmaxf(m.w, x[k].w)};
    if (C.pad < 0)
      FMaxN[(int64_t)C.v * L4 + j] = m;
    else
      part[(int64_t)C.pad * L4 + j] = m;
  }
}

__global__ __launch_bounds__(256) void k_fix_multi(int nms, const Split* __restrict__ ms, const f4v* __restrict__ part,
                                                   int Qs, f4v* __restrict__ FMaxN) {
  const int lane = threadIdx.x & 63;
  const int L4 = Qs / 4, G = 64 / L4;
  const int g = lane / L4, j = lane - g * L4;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (w >= nms) return;
  const Split M = ms[w];
  f4v m0 = {-INFINITY, -INFINITY, -INFINITY, -INFINITY}, m1 = m0;
  if (g < G) {
    int c = g;
    for (; c + G < M.count; c += 2 * G) {
      const f4v a = part[(int64_t)(M.first + c) * L4 + j], b = part[(int64_t)(M.first + c + G) * L4 + j];
      m0 = f4v{fmaxf(m0.x, a.x), fmaxf(m0.y, a.y), fmaxf(m0.z, a.z), fmaxf(m0.w, a.w)};
      m1 = f4v{fmaxf(m1.x, b.x), fmaxf(m1.y, b.y), fmaxf(m1.z, b.z), fmaxf(m1.w, b.w)};
    }
    if (c < M.count) {
      const f4v a = part[(int64_t)(M.first + c) * L4 + j];
      m0 = f4v{fmaxf(m0.x, a.x), fmaxf(m0.y, a.y), fmaxf(m0.z, a.z), fmaxf(m0.w, a.w)};
    }
  }
  f4v m = {fmaxf(m0.x, m1.x), fmaxf(m0.y, m1.y), fmaxf(m0.z, m1.z), fmaxf(m0.w, m1.w)};
  // Across the groups: lane (0, j) gathers lanes (gg, j).
  for (int gg = 1; gg < G; ++gg) {
    const int src = gg * L4 + j;
    const f4v o = {__shfl(m.x, src, 64), __shfl(m.y, src, 64), __shfl(m.z, src, 64), __shfl(m.w, src, 64)};
    m = f4v{fmaxf(m.x, o.x), fmaxf(m.y, o.y), fmaxf(m.z, o.z), fmaxf(m.w, o.w)};
  }
  if (g == 0) FMaxN[(int64_t)M.v * L4 + j] = m;
}

void launch_walk_max_fixup32(int nchunk, const Split* chunks, int nms, const Split* ms, const double* carry, int Qs,
                             float* part, float* FMaxN, hipStream_t st) {
  if (nchunk == 0) return;
  const int G = 64 / (Qs / 4);
  const int blocks = (int)std::min<int64_t>(((int64_t)nchunk + 4 * G - 1) / (4 * G), 8192);
  hipLaunchKernelGGL(k_fix_chunks, dim3(blocks), dim3(256), 0, st, nchunk, chunks, reinterpret_cast<const f4v*>(carry),
                     Qs, reinterpret_cast<f4v*>(part), reinterpret_cast<f4v*>(FMaxN));
  if (nms > 0)
    hipLaunchKernelGGL(k_fix_multi, dim3((nms + 3) / 4), dim3(256), 0, st, nms, ms, reinterpret_cast<const f4v*>(part),
                       Qs, reinterpret_cast<f4v*>(FMaxN));
}

void launch_walk_max_fixup(int nsplit, const Split* splits, const double* carry, int Q, int Qs, const QStat* qs,
                           float* FMaxN, hipStream_t st, const int* carg, int* FA) {
  if (nsplit == 0) return;
  hipLaunchKernelGGL(k_walk_max_fixup, dim3((unsigned)((nsplit + 3) / 4)), dim3(256), 0, st, nsplit, splits, carry,
                     Q, Qs, qs, FMaxN, carg, FA);
}

// ---------------------------------------------------------------------------------------------------------------
// Vertex passes.  A block takes tiles of VT (<= 64) consecutive vertices.  The tile's rows of FMaxN and VFm (and,
// in mode A, of minCEo and lsO), contiguous in HBM, and of CV are staged in LDS as FP32, into rows padded to an odd
// stride; the next tile's rows are loaded into registers (coalesced float4 loads) while the current tile is
// computed, so the loads of one tile overlap the arithmetic of the previous one.
//   p-phase : wave w handles pattern vertices p = w, w + 4, ...; lane r is tile row r, so the pattern's incidence
//             structure is wave-uniform (no divergence).  MU_pre = -CV + sum over the incidences of p in ascending q
//             (FMax term before RMax term; src/ac.hxx:164-170 made deterministic), FMax = max(fillF, FMaxN),
//             RMax = max(fillR, VFm - LSE_RE).  Mode A keeps MU = MU_pre - LSE_MU in LDS (FP64), mode OUT writes it
//             in dn x pn order, mode C reduces the column's log-sum-exp across the wave.
//   q-phase : (mode A) a thread per (q, row): VF = MU[pd] - FMax, VR = MU[ps] - RMax (src/main.cu:165-176), stored
//             with coalesced writes, and the per-q maxima / edge log-sum-exps accumulated in registers.

constexpr int kVTMax = 64;
// Staged tile limits: VT * Qs <= 2048 and VT * P <= 2048 floats, for every block size NT (256 or 512).
constexpr int kTileFloats = 2048;

// One incidence of a pattern vertex in the MU sums: its term is fmax(fill, guard(x - m - l)) with x the staged value
// at column col of the adjacent FMaxN | VFm tiles.  An FMax term has m = l = 0 (x - 0 - 0 = x exactly, -inf stays
// -inf), so both kinds take the same branch-free sequence.
struct IncRec {
  double fill, m, l;
  int col, pad;
};

static size_t vertex_lds_bytes(int mode, int VT, int P, int Q, int Qs) {
  size_t b = (size_t)(2 * P) * sizeof(double);                               // MU normalisers
  b += (size_t)(2 * Q + 1) * sizeof(IncRec);                                 // incidence records (+ one pad)
  if (mode != kModeC) b += (size_t)(P + 1) * VT * sizeof(double);            // MU
  if (mode == kModeC) b += (size_t)P * sizeof(Lse);                          // column accumulators
  const int nst = mode == kModeA ? 4 : 2;                                    // staged [VT][Qs] arrays
  b += (size_t)VT * (nst * (Qs + 1) + (P + 1)) * sizeof(float);              // staged rows
  if (mode == kModeOut) b += (size_t)VT * sizeof(int);                       // the tile's caller ids
  return (b + 15) / 16 * 16;
}

int vertex_tile(int P, int Q) {
  const int Qs = std::max(4, (Q + 3) / 4 * 4);
  int vt = kVTMax;
  if (const char* e = std::getenv("AC_VERTEX_TILE")) {  // tuning override: rows per tile, 8, 16, 32 or 64
    const int v = std::atoi(e);
    if (v == 8 || v == 16 || v == 32 || v == 64) vt = v;
  }
  while (vt > 4 && (vertex_lds_bytes(kModeA, vt, P, Q, Qs) > 65536 || vt * Qs > kTileFloats ||
                    vt * P > kTileFloats))
    vt /= 2;
  return vt;
}

// log-sum-exp of one value per lane over the wave (-inf lanes contribute nothing); fixed butterfly order.
__device__ __forceinline__ Lse wave_lse(double x) {
  double m = x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (m == -kInf) return lse_empty();
  const unsigned long long at_max = __ballot(x == m);
  const int first = __ffsll((long long)at_max) - 1;
  double e = ((int)(threadIdx.x & 63) == first) ? 0.0 : (double)__expf((float)(x - m));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
  return Lse{m, e};
}

// Butterfly merge of one log-sum-exp accumulator per lane over the wave (fixed order; lane 0's result is used).
__device__ __forceinline__ Lse wave_lse_merge(Lse s) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const Lse w{__shfl_xor(s.m, o, 64), __shfl_xor(s.t, o, 64)};
    s = lse_merge(s, w);
  }
  return s;
}

// Read-only inputs seen through the constant address space: loads at wave-uniform indexes become scalar loads, and
// branches on their values scalar branches.
typedef const __attribute__((address_space(4))) int cint_t;

// The isolated vertices' share, in closed form (block 0 of a vertex pass writes it as one extra partial).  Their
// FMaxN = VFm = -inf rows make MU_pre = -CV + K_p, K_p the sum of the fills of p's incidences (in the vertex pass's
// own operation sequence), so MU is largest at the least-cost isolated vertex and the column log-sum-exp is
// LSE(-CV) + K_p.  They have no out-edges: no edge terms.
__device__ double iso_kp(const VertexArgs& a, int p, bool sums, double acc) {
  if (sums)
    for (int k = a.inc_off[p]; k < a.inc_off[p + 1]; ++k) {
      const int code = a.inc[k], qq = code >> 1;
      acc += (code & 1) ? fmax(a.qs[qq].fillR, -kInf) : fmax(a.qs[qq].fillF, (double)(-INFINITY));
    }
  return acc;
}

__device__ QPart iso_qpart(const VertexArgs& a, int q) {
  QPart w{-kInf, -kInf, lse_empty(), lse_empty(), -1, -1};
  if (!a.iso) return w;
  const int pd = a.pd[q], ps = a.ps[q];
  const double fm = fmax(a.qs[q].fillF, (double)(-INFINITY)), rm = fmax(a.qs[q].fillR, -kInf);
  if (a.iso[pd].vmin >= 0) {
    const double mu = (iso_kp(a, pd, !a.first, -(double)a.iso[pd].cmin) - a.pst[pd].m) - a.pst[pd].l;
    w.vfmax = mu - fm;
    if (w.vfmax > -kInf) w.avf = a.iso[pd].vmin;  // as in the vertex pass, -inf records no vertex
  }
  if (a.iso[ps].vmin >= 0) {
    const double mu = (iso_kp(a, ps, !a.first, -(double)a.iso[ps].cmin) - a.pst[ps].m) - a.pst[ps].l;
    w.vrmax = mu - rm;
    if (w.vrmax > -kInf) w.avr = a.iso[ps].vmin;
  }
  return w;
}

__device__ Lse iso_lse(const VertexArgs& a, int p) {
  if (!a.iso || a.iso[p].vmin < 0) return lse_empty();
  const double m = iso_kp(a, p, !a.first, a.iso[p].m);
  return m > -kInf ? Lse{m, a.iso[p].t} : lse_empty();
}

// NT threads per block (kVW = NT / 64 waves).  kPW > 0 (mode C with P <= kVW * kPW): each lane accumulates its rows' log-sum-exp of its wave's kPW pattern vertices
// in registers over all tiles, merged across the wave once at the end.  kPW == 0: one wave-level reduction per tile.
// kArgMax (mode A): also track the vertices of the column maxima (the adjoint's tape needs them).
// kPFC: CV values each thread prefetches per tile (VT * P <= NT * kPFC); fewer registers allow more resident waves.
template <int kMode, int kPW, bool kArgMax, int NT, int kPFC = kTileFloats / NT>
__global__ __launch_bounds__(NT) void k_vertex(VertexArgs a) {
  constexpr int kVW = NT / 64, kPF4 = kTileFloats / 4 / NT;
  constexpr bool kA = kMode == kModeA;
  constexpr bool kRows = kMode == kModeOut || kMode == kModeMU;  // MU tile staged, written back row by row
  extern __shared__ double lds[];
  const int Q = a.Q, P = a.P, Qs = a.Qs, VT = a.VT;
  const int ldq = Qs + 1, ldp = P + 1;
  double* pmu = lds;                                   // [2][P]: m, l of LSE_MU
  IncRec* recs = reinterpret_cast<IncRec*>(pmu + 2 * P);  // [2Q + 1] incidence records, inc order
  double* MUs = reinterpret_cast<double*>(recs + 2 * Q + 1);  // [VT][P + 1] (modes A, OUT, MU)
  const int ldm = P + 1;
  Lse* lacc = reinterpret_cast<Lse*>(MUs + (kA || kRows ? ldm * VT : 0));    // [P] (mode C)
  float* Fs = reinterpret_cast<float*>(lacc + (kMode == kModeC ? P : 0));    // [VT][ldq] FMaxN
  float* Vs = Fs + VT * ldq;                           // [VT][ldq] VFm
  float* Ms = Vs + VT * ldq;                           // [VT][ldq] minCEo (mode A)
  float* Ls = Ms + (kA ? VT * ldq : 0);                // [VT][ldq] lsO (mode A)
  float* Cs = Ls + (kA ? VT * ldq : 0);                // [VT][ldp] CV
  int* Prow = reinterpret_cast<int*>(Cs + VT * ldp);   // [VT] caller id of each tile row (mode OUT)
  const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const bool sums = !a.first;
  cint_t* const c_incoff = (cint_t*)a.inc_off;

  for (int k = t; k <= 2 * Q; k += NT) {
    IncRec r{-kInf, 0.0, 0.0, 0, 0};  // k = 2Q: read ahead of the last incidence, never used
    if (k < 2 * Q) {
      const int code = a.inc[k], qq = code >> 1;
      const QStat x = a.qs[qq];
      r = (code & 1) ? IncRec{x.fillR, x.mRE, x.lRE, VT * ldq + qq, 0} : IncRec{x.fillF, 0.0, 0.0, qq, 0};
    }
    recs[k] = r;
  }
  for (int i = t; i < P; i += NT) {
    if (kMode != kModeC) {
      pmu[i] = a.pst[i].m;
      pmu[P + i] = a.pst[i].l;
    } else {
      lacc[i] = lse_empty();
    }
  }

  // Block 0 also writes the isolated vertices' partial, after the blocks' own ones (slot gridDim.x).
  if (blockIdx.x == 0) {
    if (kMode == kModeA)
      for (int i = t; i < Q; i += NT) a.qpart[(int64_t)gridDim.x * Q + i] = iso_qpart(a, i);
    else if (kMode == kModeC)
      for (int i = t; i < P; i += NT) a.ppart[(int64_t)gridDim.x * P + i] = iso_lse(a, i);
  }

  const int RQ = Q > 0 ? NT / Q : 0;
  const bool qthread = t < RQ * Q;
  const int q = qthread ? t % Q : 0, rq = qthread ? t / Q : 0;
  // Read once: loads inside the tile loop would follow the loop's stores (possible aliases) and wait for every
  // outstanding load, the next tile's prefetch included.
  const int q_ps = kA && qthread ? a.ps[q] : 0, q_pd = kA && qthread ? a.pd[q] : 0;
  // The thread's pattern edge's fills and RE normaliser, in registers for the whole pass: an LDS table of them would
  // take pass A (VT = 64, 24 pattern edges) just past the 40 KB that lets four blocks share a CU.
  double fF = 0.0, fR = 0.0, mR = 0.0, lR = 0.0;
  if (kA && qthread) {
    const QStat x = a.qs[q];
    fF = x.fillF, fR = x.fillR, mR = x.mRE, lR = x.lRE;
  }
  double vfmax = -kInf, vrmax = -kInf;
  int avf = -1, avr = -1;
  Lse lre = lse_empty(), lfe = lse_empty();

  const f4v* F4 = reinterpret_cast<const f4v*>(a.FMaxN);
  const f4v* V4 = reinterpret_cast<const f4v*>(a.VFm);
  const f4v* M4 = reinterpret_cast<const f4v*>(a.minCEo);
  const f4v* L4v = reinterpret_cast<const f4v*>(a.lsO);
  const int64_t ntiles = (a.n + VT - 1) / VT;
  const int64_t L4q = Qs / 4, f_end4 = a.f_end * L4q, o_lo4 = a.o_lo * L4q, o_hi4 = a.o_hi * L4q,
                o_end4 = a.o_end * L4q;
  Lse pacc[kPW > 0 ? kPW : 1];
#pragma unroll
  for (int pi = 0; pi < (kPW > 0 ? kPW : 1); ++pi) pacc[pi] = lse_empty();
  typedef f4v PfSet[4][kPF4];  // FMaxN, VFm (and in pass A minCEo, lsO)
  PfSet pfA;
  float pcA[kPFC];
  int ppA = 0;  // mode OUT: the caller id of tile row t
  // Loads of a tile into one register set; indexes past the tile are clamped to its last element (loaded, never
  // used).
  auto prefetch = [&](int64_t tl, PfSet& pf, float (&pc)[kPFC], int& pp) __attribute__((always_inline)) {
    const int64_t v0_ = tl * VT;
    const int nv_ = (int)((a.n - v0_) < VT ? (a.n - v0_) : VT);
    const int last4_ = nv_ * Qs / 4 - 1, lastc_ = nv_ * P - 1;
    const int64_t b4_ = v0_ * Qs / 4, bc_ = v0_ * P;
#pragma unroll
    for (int k = 0; k < kPF4; ++k) {
      const int64_t g_ = b4_ + min(t + k * NT, last4_);
      const bool hasF_ = !kA || g_ < f_end4, hasO_ = !kA || g_ < o_lo4 || (g_ >= o_hi4 && g_ < o_end4);
      pf[0][k] = hasF_ ? F4[g_] : f4v{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
      pf[1][k] = hasO_ ? V4[g_] : f4v{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
      if (kA) {
        pf[2][k] = hasO_ ? M4[g_] : f4v{INFINITY, INFINITY, INFINITY, INFINITY};
        pf[3][k] = hasO_ ? L4v[g_] : f4v{0.f, 0.f, 0.f, 0.f};
      }
    }
#pragma unroll
    for (int k = 0; k < kPFC; ++k) pc[k] = a.CV[bc_ + min(t + k * NT, lastc_)];
    if (kMode == kModeOut && a.perm) pp = a.perm[v0_ + min(t, nv_ - 1)];
  };
  const int64_t G = gridDim.x;
  // One tile: stage its register set in LDS, refill the set with the block's next tile, compute.
  auto run_tile = [&](int64_t tile, PfSet& pf, float (&pc)[kPFC], int& pp) __attribute__((always_inline)) {
    const int64_t v0 = tile * VT;
    const int nv = (int)((a.n - v0) < VT ? (a.n - v0) : VT);
    __syncthreads();  // tables written (first tile) / previous tile consumed
    {
      const int cnt4 = nv * Qs / 4, cntc = nv * P;
#pragma unroll
      for (int k = 0; k < kPF4; ++k) {
        const int i = t + k * NT;
        if (i < cnt4) {
          const int e = 4 * i, r = e / Qs, c = e - r * Qs;
          const int o = r * ldq + c;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            Fs[o + u] = pf[0][k][u];
            Vs[o + u] = pf[1][k][u];
            if (kA) {
              Ms[o + u] = pf[2][k][u];
              Ls[o + u] = pf[3][k][u];
            }
          }
        }
      }
#pragma unroll
      for (int k = 0; k < kPFC; ++k) {
        const int i = t + k * NT;
        if (i < cntc) {
          const int r = i / P, c = i - r * P;
          Cs[r * ldp + c] = pc[k];
        }
      }
      if (kMode == kModeOut && a.perm && t < nv) Prow[t] = pp;
    }
    __syncthreads();
    if (tile + G < ntiles) prefetch(tile + G, pf, pc, pp);
    // p-phase: wave-uniform pattern vertex, lane = tile row.
    // MU_pre of row `r` and pattern vertex `p` (-inf for rows past the tile).
#define AC_VX_MU_PRE(X, P_, R_)                                                                              \
    {                                                                                                      \
      X = -kInf;                                                                                           \
      if ((R_) < nv) {                                                                                     \
        double acc_ = -(double)Cs[(R_) * ldp + (P_)];                                                      \
        if (sums) {                                                                                        \
          const int k1_ = c_incoff[(P_) + 1];                                                              \
          int k_ = c_incoff[P_];                                                                           \
          IncRec rc_ = recs[k_];                                                                           \
          for (; k_ < k1_; ++k_) {                                                                         \
            const IncRec nx_ = recs[k_ + 1]; /* in flight while this term is formed */                     \
            const float x_ = Fs[(R_) * ldq + rc_.col];                                                     \
            const double d_ = ((double)x_ - rc_.m) - rc_.l;                                                \
            acc_ += fmax(rc_.fill, x_ == -INFINITY ? -kInf : d_);                                          \
            rc_ = nx_;                                                                                     \
          }                                                                                                \
        }                                                                                                  \
        X = acc_;                                                                                          \
      }                                                                                                    \
    }
    if (kMode == kModeC && kPW > 0) {
#pragma unroll
      for (int pi = 0; pi < kPW; ++pi) {
        const int p = wave + kVW * pi;
        if (p < P) {
          double x;
          AC_VX_MU_PRE(x, p, lane);
          lse_add(pacc[pi], x);
        }
      }
    } else {
      for (int p = wave; p < P; p += kVW) {
        const int r = lane;
        double x;
        AC_VX_MU_PRE(x, p, r);
        if (kMode == kModeC) {
          const Lse w = wave_lse(x);
          if (lane == 0) lacc[p] = lse_merge(lacc[p], w);
        } else if (r < nv) {
          const double mu = (x - pmu[p]) - pmu[P + p];
          MUs[r * ldm + p] = mu;
        }
      }
    }
#undef AC_VX_MU_PRE
    if (kRows) {
      // The tile's rows go out whole: consecutive threads write consecutive columns of a row.
      __syncthreads();
      for (int e = t; e < nv * P; e += NT) {
        const int r = e / P, c = e - r * P;
        if (kMode == kModeMU)
          a.mu_out[(v0 + r) * P + c] = MUs[r * ldm + c];
        else
          a.out[(int64_t)(a.perm ? Prow[r] : v0 + r) * P + c] = (float)MUs[r * ldm + c];
      }
    }
    if (kA) {
      __syncthreads();
      if (qthread) {
        for (int r = rq; r < nv; r += RQ) {
          const int64_t i = (v0 + r) * Qs + q;
          const float mo = Ms[r * ldq + q];
          const float lo = Ls[r * ldq + q];
          const double fm = fmax(fF, (double)Fs[r * ldq + q]);
          const float vfm = Vs[r * ldq + q];
          const double rm = fmax(fR, vfm == -INFINITY ? -kInf : ((double)vfm - mR) - lR);
          const double vf = MUs[r * ldm + q_pd] - fm;
          const double vr = MUs[r * ldm + q_ps] - rm;
          if (kArgMax) {
            if (vf > vfmax) {  // rows ascend within a thread, so the first maximum is the lowest vertex
              vfmax = vf;
              avf = (int)(v0 + r);
            }
            if (vr > vrmax) {
              vrmax = vr;
              avr = (int)(v0 + r);
            }
          } else {
            vfmax = fmax(vfmax, vf);
            vrmax = fmax(vrmax, vr);
          }
          // A vertex without out-edges keeps VFm = -inf (as filled in the prep), is no gathered source and adds no
          // edge term.
          if (mo != INFINITY) {
            const double xx = vf - (double)mo;
            a.VR[(v0 + r) * a.VRs + q] = (float)vr;
            // The row's padding columns too (never read), so that whole 128-byte lines are written.
            if (q < a.VRs - Q) a.VR[(v0 + r) * a.VRs + Q + q] = 0.f;
            a.VFm_out[i] = (float)xx;
            lse_add_finite(lre, xx + (double)lo);
            lse_add_finite(lfe, (vr - (double)mo) + (double)lo);
          }
        }
      }
    }
  };
  int64_t tile = blockIdx.x;
  if (tile < ntiles) prefetch(tile, pfA, pcA, ppA);
  for (; tile < ntiles; tile += G) run_tile(tile, pfA, pcA, ppA);
  __syncthreads();

  // Block partials, merged in a fixed order.
  if (kMode == kModeA) {
    QPart* red = reinterpret_cast<QPart*>(lds);
    if (qthread) red[rq * Q + q] = QPart{vfmax, vrmax, lre, lfe, avf, avr};
    __syncthreads();
    if (t < Q) {
      QPart acc{-kInf, -kInf, lse_empty(), lse_empty(), -1, -1};
      for (int r = 0; r < RQ; ++r) qpart_merge(acc, red[r * Q + t]);
      a.qpart[(int64_t)blockIdx.x * Q + t] = acc;
    }
  } else if (kMode == kModeC) {
    if (kPW > 0) {
#pragma unroll
      for (int pi = 0; pi < kPW; ++pi) {
        const int p = wave + kVW * pi;
        const Lse w = wave_lse_merge(pacc[pi]);
        if (p < P && lane == 0) a.ppart[(int64_t)blockIdx.x * P + p] = w;
      }
    } else {
      for (int i = t; i < P; i += NT) a.ppart[(int64_t)blockIdx.x * P + i] = lacc[i];
    }
  }
}

// Pass A runs 256-thread blocks (its register use), the lighter modes 512-thread blocks (measured at R-MAT 22).
// Pass A tiles whose CV rows fit 4 values per thread take the variant with the shorter CV prefetch: 120 instead of
// 132 VGPRs, 4 instead of 3 resident waves per SIMD.
constexpr int kNA = 256, kNO = 512;
static bool short_cv(const VertexArgs& a) { return a.VT * a.P <= kNA * 4; }

typedef void (*VertexKernel)(VertexArgs);

// The kernel a pass launches, its block size and dynamic LDS bytes.
static VertexKernel vertex_kernel(int mode, const VertexArgs& a, int* nt, size_t* lds) {
  const size_t tiles = vertex_lds_bytes(mode, a.VT, a.P, a.Q, a.Qs);
  *nt = mode == kModeA ? kNA : kNO;
  *lds = mode == kModeA ? std::max(tiles, sizeof(QPart) * kNA) : tiles;
  switch (mode) {
    case kModeA:
      if (a.track_arg) return short_cv(a) ? k_vertex<kModeA, 0, true, kNA, 4> : k_vertex<kModeA, 0, true, kNA>;
      return short_cv(a) ? k_vertex<kModeA, 0, false, kNA, 4> : k_vertex<kModeA, 0, false, kNA>;
    case kModeC:
      if (a.P <= 4 * (kNO / 64)) return k_vertex<kModeC, 4, false, kNO>;
      if (a.P <= 8 * (kNO / 64)) return k_vertex<kModeC, 8, false, kNO>;
      return k_vertex<kModeC, 0, false, kNO>;
    case kModeMU:
      return k_vertex<kModeMU, 0, false, kNO>;
    default:
      return k_vertex<kModeOut, 0, false, kNO>;
  }
}

void launch_vertex(int mode, const VertexArgs& a, int nblocks, hipStream_t st) {
  int nt = 0;
  size_t lds = 0;
  const VertexKernel k = vertex_kernel(mode, a, &nt, &lds);
  VertexArgs arg = a;
  void* args[] = {&arg};
  (void)hipLaunchKernel(reinterpret_cast<const void*>(k), dim3(nblocks), dim3(nt), args, lds, st);
}

// Blocks of a vertex pass resident at once on the device (CUs x blocks per CU at the pass's registers and LDS).
int vertex_resident_blocks(int mode, const VertexArgs& a) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  int nt = 0;
  size_t lds = 0;
  const VertexKernel k = vertex_kernel(mode, a, &nt, &lds);
  const hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k), nt, lds);
  return e == hipSuccess ? cus * std::max(per, 1) : 0;
}

// Columns [c0, c0 + ncol) of src[n][P] into dst[n][ncol].
__global__ void k_copy_cols(int64_t n, int P, int c0, int ncol, const float* __restrict__ src, float* __restrict__ dst) {
  const int64_t total = n * ncol;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = i / ncol;
    dst[i] = src[v * P + c0 + (int)(i - v * ncol)];
  }
}

__global__ void k_gather_rows(int64_t n, int ncol, const int* __restrict__ perm, const float* __restrict__ in,
                              float* __restrict__ out) {
  const int64_t total = n * ncol;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ncol, c = i - r * ncol;
    out[i] = in[(int64_t)perm[r] * ncol + c];
  }
}

void launch_gather_rows(int64_t n, int ncol, const int* perm, const float* in, float* out, hipStream_t st) {
  if (n * ncol == 0) return;
  const int blocks = (int)std::min<int64_t>((n * ncol + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(k_gather_rows, dim3(blocks), dim3(256), 0, st, n, ncol, perm, in, out);
}

void launch_copy_cols(int64_t n, int P, int c0, int ncol, const float* src, float* dst, hipStream_t st) {
  if (n == 0 || ncol == 0) return;
  const int blocks = (int)std::min<int64_t>((n * ncol + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(k_copy_cols, dim3(blocks), dim3(256), 0, st, n, P, c0, ncol, src, dst);
}

}  // namespace ac
