// The reference's operator library (src/ac.hxx) as stand-alone stage operators on caller-owned device buffers,
// FP32 values and int32 ids, same argument meaning and pattern-major layouts.  They are not used by the fused
// matcher; they exist so each reference operator can be dropped in (and checked) on its own.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/ac_mi355x.h"
#include "ac_host.hpp"

namespace {

constexpr double kInf = __builtin_huge_val();

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

// Block-wide reductions over one row (blockDim = 256).
__device__ double block_max(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}
__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}
__device__ int64_t block_min_index(int64_t v, int64_t* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] < red[threadIdx.x + w] ? red[threadIdx.x] : red[threadIdx.x + w];
    __syncthreads();
  }
  int64_t r = red[0];
  __syncthreads();
  return r;
}

// ac::RowSoftmax2 (src/ac.hxx:77-113), max-stabilised: x <- (x - max) - log1p(sum_{k != argmax} exp(x_k - max)).
__global__ __launch_bounds__(256) void k_stage_row_lsm(int64_t n_col, float* __restrict__ x) {
  __shared__ double red[256];
  __shared__ int64_t redi[256];
  float* row = x + (int64_t)blockIdx.x * n_col;
  double m = -kInf;
  for (int64_t c = threadIdx.x; c < n_col; c += 256) m = fmax(m, (double)row[c]);
  m = block_max(m, red);
  int64_t am = n_col;
  for (int64_t c = threadIdx.x; c < n_col; c += 256)
    if ((double)row[c] == m && c < am) am = c;
  am = block_min_index(am, redi);
  double t = 0.0;
  for (int64_t c = threadIdx.x; c < n_col; c += 256)
    if (c != am) t += exp((double)row[c] - m);
  t = block_sum(t, red);
  const double l = log1p(t);
  for (int64_t c = threadIdx.x; c < n_col; c += 256) row[c] = (float)(((double)row[c] - m) - l);
}

// ac::RowMax2 (src/ac.hxx:58-75)
__global__ __launch_bounds__(256) void k_stage_row_max(int64_t n_col, const float* __restrict__ in,
                                                       float* __restrict__ out) {
  __shared__ double red[256];
  const float* row = in + (int64_t)blockIdx.x * n_col;
  double m = -kInf;
  for (int64_t c = threadIdx.x; c < n_col; c += 256) m = fmax(m, (double)row[c]);
  m = block_max(m, red);
  if (threadIdx.x == 0) out[blockIdx.x] = (float)m;
}

// Float max as integer atomics: non-negative floats order like signed ints, negative floats order inversely to
// their unsigned bit patterns, so the result is exact and independent of the order of the updates.
__device__ __forceinline__ void atomic_max_float(float* addr, float v) {
  if (!(v < 0.f) && !signbit(v))
    atomicMax(reinterpret_cast<int*>(addr), __float_as_int(v));
  else
    atomicMin(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}

// ac::EdgeMaxReduce2_t (src/ac.hxx:115-149): fill, then scatter-max.
__global__ void k_stage_fill_rows(int64_t n_row, int64_t n_col, const float* __restrict__ vymax,
                                  float* __restrict__ xmax) {
  const int64_t total = n_row * n_col;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x)
    xmax[o] = vymax[o / n_col];
}
__global__ void k_stage_scatter_max(int64_t n_col_in, int64_t n_col_out, int64_t n_row, const float* __restrict__ xe,
                                    float* __restrict__ xmax, const int* __restrict__ nodes) {
  const int64_t total = n_row * n_col_in;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = o / n_col_in, c = o - r * n_col_in;
    atomic_max_float(xmax + r * n_col_out + nodes[c], xe[o]);
  }
}

// ac::ComputeMU2_t (src/ac.hxx:151-186), summed per data vertex in ascending q, FP64, rounded once.
__global__ void k_stage_compute_mu(int64_t n_col_in, int64_t n_row_in, int64_t n_col_out, int64_t n_row_out,
                                   const float* __restrict__ cv, const float* __restrict__ fmax,
                                   const float* __restrict__ rmax, const int* __restrict__ srcs,
                                   const int* __restrict__ dsts, float* __restrict__ mu) {
  const int64_t total = n_row_out * n_col_out;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = o / n_col_out, v = o - p * n_col_out;
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
  hipLaunchKernelGGL(k_stage_row_lsm, dim3((unsigned)n_row), dim3(256), 0, as_stream(stream), n_col, x);
  return launch_status("ac_row_log_softmax");
}

extern "C" int ac_row_max(int64_t n_row, int64_t n_col, const float* in, float* out, void* stream) {
  if (n_row < 0 || n_col < 0 || n_row > INT32_MAX || (n_row > 0 && (!out || (n_col > 0 && !in))))
    return bad("ac_row_max: bad arguments");
  if (n_row == 0) return AC_OK;
  hipLaunchKernelGGL(k_stage_row_max, dim3((unsigned)n_row), dim3(256), 0, as_stream(stream), n_col, in, out);
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
  hipLaunchKernelGGL(k_stage_compute_mu, dim3(grid_for(n_row_out * n_col_out, 256)), dim3(256), 0,
                     as_stream(stream), n_col_in, n_row_in, n_col_out, n_row_out, cv, fmax, rmax, srcs, dsts, mu);
  return launch_status("ac_compute_mu");
}

extern "C" int ac_transpose(const float* in, float* out, int64_t n_rows, int64_t n_cols, void* stream) {
  if (n_rows < 0 || n_cols < 0 || (n_rows * n_cols > 0 && (!in || !out))) return bad("ac_transpose: bad arguments");
  if (n_rows * n_cols == 0) return AC_OK;
  hipLaunchKernelGGL(k_stage_transpose, dim3(grid_for(n_rows * n_cols, 256)), dim3(256), 0, as_stream(stream), in,
                     out, n_rows, n_cols);
  return launch_status("ac_transpose");
}
