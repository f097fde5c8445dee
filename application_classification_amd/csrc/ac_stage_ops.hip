// The reference's operator library (src/ac.hxx) as stand-alone stage operators on caller-owned device buffers,
// FP32 values and int32 ids, same argument meaning and pattern-major layouts.  They are not used by the fused
// matcher; they exist so each reference operator can be dropped in (and checked) on its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/ac_mi355x.h"
#include "ac_host.hpp"
#include "ac_lse.hpp"

namespace {

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

int launch_status(const char* who) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ac::set_global_error(std::string(who) + ": " + hipGetErrorString(e));
    return AC_ERR_HIP;
  }
  return AC_OK;
}

int bad(const char* msg) {
  ac::set_global_error(msg);
  return AC_ERR_ARG;
}

unsigned grid_for(int64_t total, int block) {
  int64_t g = (total + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : (g > (1 << 20) ? (1 << 20) : g));
}

// ac::cdist (src/ac.hxx:33-56)
__global__ void k_stage_cdist(int64_t n_a, int64_t n_b, int64_t dim, const float* __restrict__ a,
                              const float* __restrict__ b, float* __restrict__ out) {
  const int64_t total = n_a * n_b;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = o / n_b, i = o - j * n_b;
    const float* x = a + j * dim;
    const float* y = b + i * dim;
    double acc = 0.0;
    for (int64_t k = 0; k < dim; ++k) {
      const double d = (double)x[k] - (double)y[k];
      acc = fma(d, d, acc);
    }
    out[o] = (float)sqrt(acc);
  }
}

// ---- row log-softmax (src/ac.hxx:77-113) --------------------------------------------------------------------------
// Max-stabilised: x <- (x - m) - log1p(t), m the row maximum and t the sum of exp(x - m) over the row with one maximal
// element left out, so that log-probabilities near 0 keep relative accuracy.  Three launches, HBM-bound (one read for
// the partials, one read and one write to apply):
//   k_lsm_parts : block (row, chunk) accumulates (m, t) over its chunk of the row into parts[row * nchunk + chunk];
//   k_lsm_rows  : one block per row merges the row's partials in a fixed order into fin[row] = (m, log1p(t));
//   k_lsm_apply : block (row, chunk) writes its chunk normalised.
// The chunking depends only on (n_row, n_col), so the merge order, and with it every result, is the same from call to
// call and for the allocating and the static-scratch entry points.

constexpr int kLsmBlock = 256;
constexpr int64_t kLsmChunkMin = 8192;  // elements per chunk, at least (32 per thread)
// Scratch of one call, in ac::Lse entries: n_row * nchunk partials, then n_row finals.  The static scratch of
// ac_row_log_softmax_static holds this many (1 MiB), and both entry points size their chunking to fit it.
constexpr int64_t kLsmScratch = 1 << 16;

int64_t lsm_chunks(int64_t n_row, int64_t n_col) {
  const int64_t want = (n_col + kLsmChunkMin - 1) / kLsmChunkMin;
  const int64_t room = kLsmScratch / n_row - 1;
  return std::max<int64_t>(1, std::min(want, room));
}

__device__ __forceinline__ ac::Lse block_lse(ac::Lse s, ac::Lse* red) {
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kLsmBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = ac::lse_merge(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  const ac::Lse r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void chunk_range(int64_t n_col, int64_t nchunk, int64_t chunk, int64_t& c0, int64_t& c1) {
  const int64_t csize = (n_col + nchunk - 1) / nchunk;
  c0 = std::min(n_col, chunk * csize);
  c1 = std::min(n_col, c0 + csize);
}

__global__ __launch_bounds__(kLsmBlock) void k_lsm_parts(int64_t n_col, int64_t nchunk, const float* __restrict__ x,
                                                         ac::Lse* __restrict__ parts) {
  __shared__ ac::Lse red[kLsmBlock];
  const int64_t row = blockIdx.x / nchunk, chunk = blockIdx.x - row * nchunk;
  int64_t c0, c1;
  chunk_range(n_col, nchunk, chunk, c0, c1);
  const float* xr = x + row * n_col;
  ac::Lse s = ac::lse_empty();
  int64_t c = c0 + threadIdx.x;
  for (; c + 3 * kLsmBlock < c1; c += 4 * kLsmBlock) {
    const float a = xr[c], b = xr[c + kLsmBlock], d = xr[c + 2 * kLsmBlock], e = xr[c + 3 * kLsmBlock];
    ac::lse_add(s, (double)a);
    ac::lse_add(s, (double)b);
    ac::lse_add(s, (double)d);
    ac::lse_add(s, (double)e);
  }
  for (; c < c1; c += kLsmBlock) ac::lse_add(s, (double)xr[c]);
  s = block_lse(s, red);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

__global__ __launch_bounds__(kLsmBlock) void k_lsm_rows(int64_t nchunk, const ac::Lse* __restrict__ parts,
                                                        ac::Lse2* __restrict__ fin) {
  __shared__ ac::Lse red[kLsmBlock];
  const int64_t row = blockIdx.x;
  ac::Lse s = ac::lse_empty();
  for (int64_t k = threadIdx.x; k < nchunk; k += kLsmBlock) s = ac::lse_merge(s, parts[row * nchunk + k]);
  s = block_lse(s, red);
  if (threadIdx.x == 0) fin[row] = ac::Lse2{s.m, log1p(s.t)};
}

__global__ __launch_bounds__(kLsmBlock) void k_lsm_apply(int64_t n_col, int64_t nchunk, float* __restrict__ x,
                                                         const ac::Lse2* __restrict__ fin) {
  const int64_t row = blockIdx.x / nchunk, chunk = blockIdx.x - row * nchunk;
  int64_t c0, c1;
  chunk_range(n_col, nchunk, chunk, c0, c1);
  const ac::Lse2 f = fin[row];
  float* xr = x + row * n_col;
  for (int64_t c = c0 + threadIdx.x; c < c1; c += kLsmBlock) xr[c] = (float)(((double)xr[c] - f.m) - f.l);
}

// The three launches on `scratch`, which holds lsm_entries(n_row, n_col) entries.
int64_t lsm_entries(int64_t n_row, int64_t n_col) { return n_row * (lsm_chunks(n_row, n_col) + 1); }

int enqueue_row_lsm(int64_t n_row, int64_t n_col, float* x, ac::Lse* scratch, hipStream_t st, const char* who) {
  const int64_t nchunk = lsm_chunks(n_row, n_col);
  const int64_t nblk = n_row * nchunk;
  ac::Lse2* fin = reinterpret_cast<ac::Lse2*>(scratch + nblk);
  hipLaunchKernelGGL(k_lsm_parts, dim3((unsigned)nblk), dim3(kLsmBlock), 0, st, n_col, nchunk, x, scratch);
  hipLaunchKernelGGL(k_lsm_rows, dim3((unsigned)n_row), dim3(kLsmBlock), 0, st, nchunk, scratch, fin);
  hipLaunchKernelGGL(k_lsm_apply, dim3((unsigned)nblk), dim3(kLsmBlock), 0, st, n_col, nchunk, x, fin);
  return launch_status(who);
}

// Scratch of ac_row_log_softmax_static: one copy per GPU, held until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress returns on each device; the array's name itself is a host-side stand-in.
static __device__ __align__(256) ac::Lse scratch[kLsmScratch];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
ac::Lse* g_scratch_ptr[kMaxDevices] = {};

// The scratch's address on the current device, looked up on the first call there (with that device current).
int static_scratch(ac::Lse** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    ac::set_global_error(std::string("ac_row_log_softmax_static: hipGetDevice: ") + hipGetErrorString(e));
    return AC_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) return bad("ac_row_log_softmax_static: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ac::set_global_error(std::string("ac_row_log_softmax_static: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return AC_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<ac::Lse*>(p);
  }
  *out = g_scratch_ptr[dev];
  return AC_OK;
}

// Float max as integer atomics: non-negative floats order like signed ints, negative floats order inversely to
// their unsigned bit patterns, so the result is exact and independent of the order of the updates.
__device__ __forceinline__ void atomic_max_float(float* addr, float v) {
  if (!(v < 0.f) && !signbit(v))
    atomicMax(reinterpret_cast<int*>(addr), __float_as_int(v));
  else
    atomicMin(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}

// ---- row max (src/ac.hxx:58-75) -----------------------------------------------------------------------------------
// out[r] starts at -inf (k_stage_fill_rows without fill values), then block (row, chunk) reduces its chunk with wave
// shuffles and folds the result in with one atomic.  Max is exact and order-free, so no scratch is needed.
constexpr int kMaxBlock = 256;

__global__ __launch_bounds__(kMaxBlock) void k_stage_row_max(int64_t n_col, int64_t nchunk,
                                                             const float* __restrict__ in, float* __restrict__ out) {
  __shared__ float red[kMaxBlock / 64];
  const int64_t row = blockIdx.x / nchunk, chunk = blockIdx.x - row * nchunk;
  int64_t c0, c1;
  chunk_range(n_col, nchunk, chunk, c0, c1);
  const float* xr = in + row * n_col;
  float m = -INFINITY;
  int64_t c = c0 + threadIdx.x;
  for (; c + 3 * kMaxBlock < c1; c += 4 * kMaxBlock)
    m = fmaxf(fmaxf(m, fmaxf(xr[c], xr[c + kMaxBlock])), fmaxf(xr[c + 2 * kMaxBlock], xr[c + 3 * kMaxBlock]));
  for (; c < c1; c += kMaxBlock) m = fmaxf(m, xr[c]);
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kMaxBlock / 64; ++w) m = fmaxf(m, red[w]);
    atomic_max_float(out + row, m);
  }
}

// ac::EdgeMaxReduce2_t (src/ac.hxx:115-149): fill, then scatter-max.
__global__ void k_stage_fill_rows(int64_t n_row, int64_t n_col, const float* __restrict__ vymax,
                                  float* __restrict__ xmax) {
  const int64_t total = n_row * n_col;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x)
    xmax[o] = vymax ? vymax[o / n_col] : -INFINITY;
}
__global__ void k_stage_scatter_max(int64_t n_col_in, int64_t n_col_out, int64_t n_row, const float* __restrict__ xe,
                                    float* __restrict__ xmax, const int* __restrict__ nodes) {
  const int64_t total = n_row * n_col_in;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = o / n_col_in, c = o - r * n_col_in;
    atomic_max_float(xmax + r * n_col_out + nodes[c], xe[o]);
  }
}

// ac::ComputeMU2_t (src/ac.hxx:151-186), summed per data vertex in ascending q, FP64, rounded once.  Block (v tile,
// p): the q loop is the same for the whole block, so the endpoint tests are scalar and only the pattern edges incident
// to p issue (coalesced) vector loads.
__global__ __launch_bounds__(256) void k_stage_compute_mu(int64_t n_col_in, int64_t n_row_in, int64_t n_col_out,
                                                          int64_t n_row_out, const float* __restrict__ cv,
                                                          const float* __restrict__ fmax,
                                                          const float* __restrict__ rmax,
                                                          const int* __restrict__ srcs, const int* __restrict__ dsts,
                                                          float* __restrict__ mu) {
  for (int64_t p = blockIdx.y; p < n_row_out; p += gridDim.y)
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n_col_out;
         v += (int64_t)gridDim.x * blockDim.x) {
      const int64_t o = p * n_col_out + v;
      double acc = -(double)cv[o];
      if (v < n_col_in)
        for (int64_t q = 0; q < n_row_in; ++q) {
          if (dsts[q] == p) acc += (double)fmax[q * n_col_in + v];
          if (srcs[q] == p) acc += (double)rmax[q * n_col_in + v];
        }
      mu[o] = (float)acc;
    }
}

// ac::transpose (src/ac.hxx:17-31)
__global__ void k_stage_transpose(const float* __restrict__ in, float* __restrict__ out, int64_t n_rows,
                                  int64_t n_cols) {
  const int64_t total = n_rows * n_cols;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = o / n_rows, r = o - c * n_rows;
    out[o] = in[r * n_cols + c];
  }
}

}  // namespace

extern "C" int ac_cdist(int64_t n_a, int64_t n_b, int64_t dim, const float* feats_a, const float* feats_b, float* out,
                        void* stream) {
  if (n_a < 0 || n_b < 0 || dim < 0 || (n_a * n_b > 0 && (!out || (dim > 0 && (!feats_a || !feats_b)))))
    return bad("ac_cdist: bad arguments");
  if (n_a * n_b == 0) return AC_OK;
  hipLaunchKernelGGL(k_stage_cdist, dim3(grid_for(n_a * n_b, 256)), dim3(256), 0, as_stream(stream), n_a, n_b, dim,
                     feats_a, feats_b, out);
  return launch_status("ac_cdist");
}

extern "C" int ac_row_log_softmax(int64_t n_row, int64_t n_col, float* x, void* stream) {
  if (n_row < 0 || n_col < 0 || n_row > INT32_MAX || (n_row * n_col > 0 && !x)) return bad("ac_row_log_softmax: bad arguments");
  if (n_row * n_col == 0) return AC_OK;
  hipStream_t st = as_stream(stream);
  // Stream-ordered temporary (the reference's RowSoftmax2 allocates its own too, src/ac.hxx:108-113).
  ac::Lse* tmp = nullptr;
  hipError_t e = hipMallocAsync((void**)&tmp, (size_t)lsm_entries(n_row, n_col) * sizeof(ac::Lse), st);
  if (e != hipSuccess) {
    ac::set_global_error(std::string("ac_row_log_softmax: hipMallocAsync: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? AC_ERR_OOM : AC_ERR_HIP;
  }
  const int rc = enqueue_row_lsm(n_row, n_col, x, tmp, st, "ac_row_log_softmax");
  e = hipFreeAsync(tmp, st);
  if (rc != AC_OK) return rc;
  if (e != hipSuccess) {
    ac::set_global_error(std::string("ac_row_log_softmax: hipFreeAsync: ") + hipGetErrorString(e));
    return AC_ERR_HIP;
  }
  return AC_OK;
}

extern "C" int ac_row_log_softmax_static(int64_t n_row, int64_t n_col, float* x, void* stream) {
  if (n_row < 0 || n_col < 0 || (n_row * n_col > 0 && !x)) return bad("ac_row_log_softmax_static: bad arguments");
  if (n_row > kLsmScratch / 2)
    return bad("ac_row_log_softmax_static: more than 32768 rows (its fixed scratch holds 65536 partials)");
  if (n_row * n_col == 0) return AC_OK;
  ac::Lse* scr = nullptr;
  const int rc = static_scratch(&scr);
  if (rc != AC_OK) return rc;
  return enqueue_row_lsm(n_row, n_col, x, scr, as_stream(stream), "ac_row_log_softmax_static");
}

extern "C" int ac_row_max(int64_t n_row, int64_t n_col, const float* in, float* out, void* stream) {
  if (n_row < 0 || n_col < 0 || n_row > INT32_MAX || (n_row > 0 && (!out || (n_col > 0 && !in))))
    return bad("ac_row_max: bad arguments");
  if (n_row == 0) return AC_OK;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_stage_fill_rows, dim3(grid_for(n_row, 256)), dim3(256), 0, st, n_row, (int64_t)1,
                     (const float*)nullptr, out);
  if (n_col > 0) {
    const int64_t nchunk = std::max<int64_t>(
        1, std::min<int64_t>((n_col + kLsmChunkMin - 1) / kLsmChunkMin, (int64_t)INT32_MAX / n_row));
    hipLaunchKernelGGL(k_stage_row_max, dim3((unsigned)(n_row * nchunk)), dim3(kMaxBlock), 0, st, n_col, nchunk, in,
                       out);
  }
  return launch_status("ac_row_max");
}

extern "C" int ac_edge_max_reduce(int64_t n_col_in, int64_t n_col_out, int64_t n_row, const float* vymax,
                                  const float* xe, float* xmax, const int32_t* nodes, void* stream) {
  if (n_col_in < 0 || n_col_out < 0 || n_row < 0 || (n_row * n_col_out > 0 && (!vymax || !xmax)) ||
      (n_row * n_col_in > 0 && (!xe || !nodes)))
    return bad("ac_edge_max_reduce: bad arguments");
  hipStream_t st = as_stream(stream);
  if (n_row * n_col_out > 0)
    hipLaunchKernelGGL(k_stage_fill_rows, dim3(grid_for(n_row * n_col_out, 256)), dim3(256), 0, st, n_row,
                       n_col_out, vymax, xmax);
  if (n_row * n_col_in > 0)
    hipLaunchKernelGGL(k_stage_scatter_max, dim3(grid_for(n_row * n_col_in, 256)), dim3(256), 0, st, n_col_in,
                       n_col_out, n_row, xe, xmax, nodes);
  return launch_status("ac_edge_max_reduce");
}

extern "C" int ac_compute_mu(int64_t n_col_in, int64_t n_row_in, int64_t n_col_out, int64_t n_row_out,
                             const float* cv, const float* fmax, const float* rmax, const int32_t* srcs,
                             const int32_t* dsts, float* mu, void* stream) {
  if (n_col_in < 0 || n_row_in < 0 || n_col_out < 0 || n_row_out < 0 ||
      (n_row_out * n_col_out > 0 && (!cv || !mu)) || (n_row_in > 0 && (!fmax || !rmax || !srcs || !dsts)))
    return bad("ac_compute_mu: bad arguments");
  if (n_row_out * n_col_out == 0) return AC_OK;
  const dim3 grid((unsigned)std::min<int64_t>((n_col_out + 255) / 256, 1 << 16),
                  (unsigned)std::min<int64_t>(n_row_out, 4096));
  hipLaunchKernelGGL(k_stage_compute_mu, grid, dim3(256), 0, as_stream(stream), n_col_in, n_row_in, n_col_out,
                     n_row_out, cv, fmax, rmax, srcs, dsts, mu);
  return launch_status("ac_compute_mu");
}

extern "C" int ac_transpose(const float* in, float* out, int64_t n_rows, int64_t n_cols, void* stream) {
  if (n_rows < 0 || n_cols < 0 || (n_rows * n_cols > 0 && (!in || !out))) return bad("ac_transpose: bad arguments");
  if (n_rows * n_cols == 0) return AC_OK;
  hipLaunchKernelGGL(k_stage_transpose, dim3(grid_for(n_rows * n_cols, 256)), dim3(256), 0, as_stream(stream), in,
                     out, n_rows, n_cols);
  return launch_status("ac_transpose");
}
