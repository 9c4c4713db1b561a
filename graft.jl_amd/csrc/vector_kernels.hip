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

// A non-null gate with gate[0] == 0 turns the pass off: zero partials, no
// vector traffic (a capturable graph keeps the launch and skips the work).
__device__ __forceinline__ bool gated_off(const double* gate) { return gate != nullptr && gate[0] == 0.0; }

__global__ void __launch_bounds__(kVecThreads)
    dotc_partials_kernel(int64_t n, const double2* __restrict__ x, const double2* __restrict__ y,
                         double* __restrict__ partials, const double* __restrict__ gate) {
  __shared__ double sh[2 * (kVecThreads / 64)];
  if (gated_off(gate)) {
    if (threadIdx.x < 2) partials[2 * blockIdx.x + threadIdx.x] = 0.0;
    return;
  }
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

// Multi-dot <V_j, y> in passes of CH vectors over y: the accumulators keep
// compile-time indices (registers, no scratch); the wave sums go through
// LDS in a fixed order.  partials layout: [block][j][re, im].
template <int CH>
__global__ void __launch_bounds__(kVecThreads)
    multi_dotc_partials_kernel(int64_t n, int k, PtrList V, const double2* __restrict__ y,
                               double* __restrict__ partials, const double* __restrict__ gate) {
  constexpr int kWaves = kVecThreads / 64;
  __shared__ double sh[kWaves][2 * CH];
  if (gated_off(gate)) {
    for (int t = threadIdx.x; t < 2 * k; t += blockDim.x) partials[(size_t)blockIdx.x * 2 * k + t] = 0.0;
    return;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int j0 = 0; j0 < k; j0 += CH) {
    const int kc = min(CH, k - j0);
    double re[CH], im[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      re[j] = 0.0;
      im[j] = 0.0;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
      const double2 b = y[i];
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        if (j < kc) {
          const double2 a = V.p[j0 + j][i];
          re[j] += a.x * b.x + a.y * b.y;   // conj(a) * b
          im[j] += a.x * b.y - a.y * b.x;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      if (j < kc) {
        double r = re[j], m = im[j];
        for (int off = 32; off > 0; off >>= 1) {
          r += __shfl_down(r, off, 64);
          m += __shfl_down(m, off, 64);
        }
        if (lane == 0) {
          sh[wave][2 * j] = r;
          sh[wave][2 * j + 1] = m;
        }
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * kc; t += blockDim.x) {
      double acc = 0.0;
      for (int w = 0; w < kWaves; ++w) acc += sh[w][t];
      partials[(size_t)blockIdx.x * 2 * k + 2 * j0 + t] = acc;
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

// y += sign * sum_j c_j V_j with the coefficients read from device memory
// (so that a captured graph sees the values of the replay, not of capture).
template <int KMAX>
__global__ void __launch_bounds__(kVecThreads)
    multi_axpy_kernel(int64_t n, int k, PtrList V, const double* __restrict__ c, double sign,
                      double2* __restrict__ y, const double* __restrict__ gate) {
  if (gated_off(gate)) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double2 acc = y[i];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      if (j < k) {
        const double2 v = V.p[j][i];
        const double cr = sign * c[2 * j], ci = sign * c[2 * j + 1];
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
                                int nblocks, hipStream_t stream, const double* gate) {
  hipLaunchKernelGGL(dotc_partials_kernel, dim3(nblocks), dim3(kVecThreads), 0, stream, n,
                     (const double2*)x, (const double2*)y, partials, gate);
  return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* partials, int nblocks, int k, double* out,
                                  hipStream_t stream) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kVecThreads), 0, stream, partials,
                     nblocks, k, out);
  return hipGetLastError();
}

hipError_t launch_multi_dotc_partials(int64_t n, int k, const PtrList& V, const void* y,
                                      double* partials, int nblocks, hipStream_t stream,
                                      const double* gate) {
  if (k <= 8)
    hipLaunchKernelGGL(multi_dotc_partials_kernel<8>, dim3(nblocks), dim3(kVecThreads), 0, stream,
                       n, k, V, (const double2*)y, partials, gate);
  else
    hipLaunchKernelGGL(multi_dotc_partials_kernel<16>, dim3(nblocks), dim3(kVecThreads), 0, stream,
                       n, k, V, (const double2*)y, partials, gate);
  return hipGetLastError();
}

hipError_t launch_axpby(int64_t n, double ar, double ai, const void* x, double br, double bi,
                        void* y, hipStream_t stream) {
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(kVecThreads), 0, stream, n, ar, ai,
                     (const double2*)x, br, bi, (double2*)y);
  return hipGetLastError();
}

hipError_t launch_multi_axpy(int64_t n, int k, const PtrList& V, const double* c, double sign,
                             void* y, hipStream_t stream, const double* gate) {
  if (k <= 8) {
    hipLaunchKernelGGL(multi_axpy_kernel<8>, dim3(grid_for(n)), dim3(kVecThreads), 0, stream, n, k,
                       V, c, sign, (double2*)y, gate);
  } else if (k <= 32) {
    hipLaunchKernelGGL(multi_axpy_kernel<32>, dim3(grid_for(n)), dim3(kVecThreads), 0, stream, n,
                       k, V, c, sign, (double2*)y, gate);
  } else {
    hipLaunchKernelGGL(multi_axpy_kernel<kMaxMultiVec>, dim3(grid_for(n)), dim3(kVecThreads), 0,
                       stream, n, k, V, c, sign, (double2*)y, gate);
  }
  return hipGetLastError();
}

// out = sqrt(Re <x, x>) from the reduced dot product, and x /= out.
__global__ void sqrt_real_kernel(const double* __restrict__ dot2, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = sqrt(fmax(dot2[0], 0.0));
}

__global__ void __launch_bounds__(kVecThreads)
    scale_inv_kernel(int64_t n, const double* __restrict__ beta, double2* __restrict__ x) {
  const double s = 1.0 / beta[0];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double2 v = x[i];
    v.x *= s;
    v.y *= s;
    x[i] = v;
  }
}

// After the first Gram-Schmidt pass: beta = ||w_1|| and gate = 1 when the
// pass removed more than 1 - eta of the norm (||w_1||^2 < eta^2 ||w_0||^2
// with ||w_0||^2 = ||w_1||^2 + sum |c_j|^2 for an orthonormal basis).
__global__ void reorth_gate_kernel(const double* __restrict__ c, int k,
                                   const double* __restrict__ dot2, double eta,
                                   double* __restrict__ beta, double* __restrict__ gate) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += c[2 * j] * c[2 * j] + c[2 * j + 1] * c[2 * j + 1];
    const double n2 = fmax(dot2[0], 0.0);
    beta[0] = sqrt(n2);
    gate[0] = (n2 < eta * eta * (n2 + s)) ? 1.0 : 0.0;
  }
}

// After the (possibly skipped) second pass: beta = ||w_2|| when it ran.
__global__ void reorth_finish_kernel(const double* __restrict__ dot2,
                                     const double* __restrict__ gate, double* __restrict__ beta) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && gate[0] != 0.0) beta[0] = sqrt(fmax(dot2[0], 0.0));
}

hipError_t launch_reorth_gate(const double* c, int k, const double* dot2, double eta, double* beta,
                              double* gate, hipStream_t stream) {
  hipLaunchKernelGGL(reorth_gate_kernel, dim3(1), dim3(64), 0, stream, c, k, dot2, eta, beta, gate);
  return hipGetLastError();
}

hipError_t launch_reorth_finish(const double* dot2, const double* gate, double* beta,
                                hipStream_t stream) {
  hipLaunchKernelGGL(reorth_finish_kernel, dim3(1), dim3(64), 0, stream, dot2, gate, beta);
  return hipGetLastError();
}

hipError_t launch_sqrt_real(const double* dot2, double* out, hipStream_t stream) {
  hipLaunchKernelGGL(sqrt_real_kernel, dim3(1), dim3(64), 0, stream, dot2, out);
  return hipGetLastError();
}

hipError_t launch_scale_inv(int64_t n, const double* beta, void* x, hipStream_t stream) {
  hipLaunchKernelGGL(scale_inv_kernel, dim3(grid_for(n)), dim3(kVecThreads), 0, stream, n, beta,
                     (double2*)x);
  return hipGetLastError();
}

// O_j = sum_l c[l + j k] V_l for j < p (column-major k x p coefficients).
// Outputs may alias inputs: the thread that owns element i loads V_l[i] for
// every l before it stores any O_j[i], and no other thread touches index i.
template <int KMAX>
__global__ void __launch_bounds__(kVecThreads)
    block_combine_kernel(int64_t n, int k, PtrList V, int p, PtrList O,
                         const double2* __restrict__ c) {
  __shared__ double2 cs[kMaxMultiVec * kMaxCombine];
  for (int t = threadIdx.x; t < k * p; t += blockDim.x) cs[t] = c[t];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double2 x[KMAX];
#pragma unroll
    for (int l = 0; l < KMAX; ++l)
      if (l < k) x[l] = V.p[l][i];
    for (int j = 0; j < p; ++j) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int l = 0; l < KMAX; ++l) {
        if (l < k) {
          const double2 cc = cs[l + j * k];
          acc.x += cc.x * x[l].x - cc.y * x[l].y;
          acc.y += cc.x * x[l].y + cc.y * x[l].x;
        }
      }
      const_cast<double2*>(O.p[j])[i] = acc;
    }
  }
}

hipError_t launch_block_combine(int64_t n, int k, const PtrList& V, int p, const PtrList& O,
                                const double* c, hipStream_t stream) {
  const double2* cc = (const double2*)c;
  if (k <= 8) {
    hipLaunchKernelGGL(block_combine_kernel<8>, dim3(grid_for(n)), dim3(kVecThreads), 0, stream, n,
                       k, V, p, O, cc);
  } else if (k <= 16) {
    hipLaunchKernelGGL(block_combine_kernel<16>, dim3(grid_for(n)), dim3(kVecThreads), 0, stream,
                       n, k, V, p, O, cc);
  } else if (k <= 32) {
    hipLaunchKernelGGL(block_combine_kernel<32>, dim3(grid_for(n)), dim3(kVecThreads), 0, stream,
                       n, k, V, p, O, cc);
  } else {
    hipLaunchKernelGGL(block_combine_kernel<kMaxMultiVec>, dim3(grid_for(n)), dim3(kVecThreads), 0,
                       stream, n, k, V, p, O, cc);
  }
  return hipGetLastError();
}

}  // namespace graft
