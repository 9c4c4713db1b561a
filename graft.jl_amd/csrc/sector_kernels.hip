// Copies between block-sparse storage and fused per-charge matrices.
//
// The fused sector path runs one contraction step on U(1) x U(1) block-sparse
// tensors as (1) gather of every operand block into the dense matrix of its
// charge component (rows: open-leg sector tuples, columns: contracted sector
// tuples), (2) one grouped dense GEMM over the components, (3) scatter of the
// result matrices into the output blocks.  Both copies walk the blocked
// tensor's elements in storage order (one thread per element, coalesced on
// the blocked side); a chunk table narrows the binary search for the block
// that owns an element to the blocks that start inside its chunk.
#include "heff_internal.h"

namespace graft {

static constexpr int kPermThreads = 256;

__global__ void __launch_bounds__(kPermThreads)
    perm_copy_kernel(int to_fused, const double2* __restrict__ src, double2* __restrict__ dst,
                     const PermBlock* __restrict__ blocks, int nblocks,
                     const int32_t* __restrict__ chunk_first, int64_t total) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t c = e / kPermChunk;
    int lo = chunk_first[c], hi = chunk_first[c + 1];
    // last block with base <= e in [lo, hi]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (blocks[mid].base <= e) lo = mid;
      else hi = mid - 1;
    }
    const PermBlock& b = blocks[lo];
    int64_t local = e - b.base;
    if (local < 0 || local >= b.size) continue;     // a gap between blocks
    if (b.other < 0) {
      if (!to_fused) dst[e] = double2{0.0, 0.0};
      continue;
    }
    // blocks hold < 2^31 elements: 32-bit index arithmetic
    int64_t o = b.other;
    uint32_t loc = (uint32_t)local;
    for (int l = 0; l < b.rank; ++l) {
      const uint32_t d = (uint32_t)b.dims[l];
      const uint32_t q = loc / d;
      o += (int64_t)(loc - q * d) * b.ostr[l];
      loc = q;
    }
    if (to_fused) dst[o] = src[e];
    else dst[e] = src[o];
  }
}

__global__ void __launch_bounds__(kPermThreads) fill_zero_kernel(double2* __restrict__ x, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] = double2{0.0, 0.0};
}

static int perm_grid(int64_t n) {
  int64_t g = (n + kPermThreads - 1) / kPermThreads;
  if (g > 16384) g = 16384;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_perm_copy(bool to_fused, const void* blocked, void* fused, const PermBlock* blocks,
                            int nblocks, const int32_t* chunk_first, int64_t total,
                            hipStream_t stream) {
  if (total <= 0 || nblocks <= 0) return hipSuccess;
  const double2* src = to_fused ? (const double2*)blocked : (const double2*)fused;
  double2* dst = to_fused ? (double2*)fused : (double2*)blocked;
  hipLaunchKernelGGL(perm_copy_kernel, dim3(perm_grid(total)), dim3(kPermThreads), 0, stream,
                     to_fused ? 1 : 0, src, dst, blocks, nblocks, chunk_first, total);
  return hipGetLastError();
}

hipError_t launch_fill_zero(void* x, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_zero_kernel, dim3(perm_grid(n)), dim3(kPermThreads), 0, stream,
                     (double2*)x, n);
  return hipGetLastError();
}

}  // namespace graft
