// Krylov vector operations on complex128 vectors (HBM-bound).
//
// Reductions are deterministic: a fixed grid of kReduceBlocks blocks writes
// one partial per block, and a single block sums the partials in a fixed
// order, so repeated calls on the same data give bit-identical results.
#include "heff_internal.h"

namespace graft {

static constexpr int kVecThreads = 256;

__device__ __forceinline__ void block_sum2(double& re, double& im, double* sh) {
  // wave reduction then cross-wave via LDS (fixed order)
  for (int off = 32; off > 0; off >>= 1) {
    re += __shfl_down(re, off, 64);
    im += __shfl_down(im, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh[2 * wave] = re;
    sh[2 * wave + 1] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0, i = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      r += sh[2 * w];
      i += sh[2 * w + 1];
    }
    re = r;
    im = i;
  }
}

__global__ void __launch_bounds__(kVecThreads)
    dotc_partials_kernel(int64_t n, const double2* __restrict__ x, const double2* __restrict__ y,
                         double* __restrict__ partials) {
  __shared__ double sh[2 * (kVecThreads / 64)];
  double re = 0.0, im = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2 a = x[i], b = y[i];
    re += a.x * b.x + a.y * b.y;  // conj(a) * b
    im += a.x * b.y - a.y * b.x;
  }
  block_sum2(re, im, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = re;
    partials[2 * blockIdx.x + 1] = im;
  }
}

// partials layout: [block][j][re, im] for k values per block.
__global__ void __launch_bounds__(kVecThreads)
    reduce_partials_kernel(const double* __restrict__ partials, int nblocks, int k,
                           double* __restrict__ out) {
  __shared__ double sh[2 * (kVecThreads / 64)];
  for (int j = 0; j < k; ++j) {
    double re = 0.0, im = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
      re += partials[(size_t)b * 2 * k + 2 * j];
      im += partials[(size_t)b * 2 * k + 2 * j + 1];
    }
    block_sum2(re, im, sh);
    if (threadIdx.x == 0) {
      out[2 * j] = re;
      out[2 * j + 1] = im;
    }
    __syncthreads();
  }
}

template <int KMAX>
__global__ void __launch_bounds__(kVecThreads)
    multi_dotc_partials_kernel(int64_t n, int k, const double2* const* __restrict__ V,
                               const double2* __restrict__ y, double* __restrict__ partials) {
  __shared__ double sh[2 * (kVecThreads / 64)];
  double re[KMAX], im[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) { re[j] = 0.0; im[j] = 0.0; }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2 b = y[i];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      if (j < k) {
        const double2 a = V[j][i];
        re[j] += a.x * b.x + a.y * b.y;
        im[j] += a.x * b.y - a.y * b.x;
      }
    }
  }
  for (int j = 0; j < k; ++j) {
    double r = re[j], m = im[j];
    block_sum2(r, m, sh);
    if (threadIdx.x == 0) {
      partials[(size_t)blockIdx.x * 2 * k + 2 * j] = r;
      partials[(size_t)blockIdx.x * 2 * k + 2 * j + 1] = m;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kVecThreads)
    axpby_kernel(int64_t n, double ar, double ai, const double2* __restrict__ x, double br,
                 double bi, double2* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2 a = x[i], b = y[i];
    double2 r;
    r.x = ar * a.x - ai * a.y + br * b.x - bi * b.y;
    r.y = ar * a.y + ai * a.x + br * b.y + bi * b.x;
    y[i] = r;
  }
}

template <int KMAX>
__global__ void __launch_bounds__(kVecThreads)
    multi_axpy_kernel(int64_t n, int k, const double2* const* __restrict__ V,
                      const double* __restrict__ c, double2* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double2 acc = y[i];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      if (j < k) {
        const double2 v = V[j][i];
        const double cr = c[2 * j], ci = c[2 * j + 1];
        acc.x += cr * v.x - ci * v.y;
        acc.y += cr * v.y + ci * v.x;
      }
    }
    y[i] = acc;
  }
}

static int grid_for(int64_t n) {
  int64_t g = (n + kVecThreads - 1) / kVecThreads;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_dotc_partials(int64_t n, const void* x, const void* y, double* partials,
                                int nblocks, hipStream_t stream) {
  hipLaunchKernelGGL(dotc_partials_kernel, dim3(nblocks), dim3(kVecThreads), 0, stream, n,
                     (const double2*)x, (const double2*)y, partials);
  return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* partials, int nblocks, int k, double* out,
                                  hipStream_t stream) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kVecThreads), 0, stream, partials,
                     nblocks, k, out);
  return hipGetLastError();
}

hipError_t launch_multi_dotc_partials(int64_t n, int k, const void* const* V, const void* y,
                                      double* partials, int nblocks, hipStream_t stream) {
  if (k <= 8) {
    hipLaunchKernelGGL(multi_dotc_partials_kernel<8>, dim3(nblocks), dim3(kVecThreads), 0, stream,
                       n, k, (const double2* const*)V, (const double2*)y, partials);
  } else if (k <= 32) {
    hipLaunchKernelGGL(multi_dotc_partials_kernel<32>, dim3(nblocks), dim3(kVecThreads), 0,
                       stream, n, k, (const double2* const*)V, (const double2*)y, partials);
  } else {
    hipLaunchKernelGGL(multi_dotc_partials_kernel<kMaxMultiVec>, dim3(nblocks), dim3(kVecThreads),
                       0, stream, n, k, (const double2* const*)V, (const double2*)y, partials);
  }
  return hipGetLastError();
}

hipError_t launch_axpby(int64_t n, double ar, double ai, const void* x, double br, double bi,
                        void* y, hipStream_t stream) {
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(kVecThreads), 0, stream, n, ar, ai,
                     (const double2*)x, br, bi, (double2*)y);
  return hipGetLastError();
}

hipError_t launch_multi_axpy(int64_t n, int k, const void* const* V, const double* c, void* y,
                             hipStream_t stream) {
  if (k <= 8) {
    hipLaunchKernelGGL(multi_axpy_kernel<8>, dim3(grid_for(n)), dim3(kVecThreads), 0, stream, n, k,
                       (const double2* const*)V, c, (double2*)y);
  } else if (k <= 32) {
    hipLaunchKernelGGL(multi_axpy_kernel<32>, dim3(grid_for(n)), dim3(kVecThreads), 0, stream, n,
                       k, (const double2* const*)V, c, (double2*)y);
  } else {
    hipLaunchKernelGGL(multi_axpy_kernel<kMaxMultiVec>, dim3(grid_for(n)), dim3(kVecThreads), 0,
                       stream, n, k, (const double2* const*)V, c, (double2*)y);
  }
  return hipGetLastError();
}

}  // namespace graft
