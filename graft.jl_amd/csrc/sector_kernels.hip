// Copies between block-sparse storage and fused per-charge matrices.
//
// The fused sector path runs one contraction step on U(1) x U(1) block-sparse
// tensors as (1) gather of every operand block into the dense matrix of its
// charge component (rows: open-leg sector tuples, columns: contracted sector
// tuples), (2) one grouped dense GEMM over the components, (3) scatter of the
// result matrices into the output blocks.  Once per plan, a map kernel walks
// each blocked tensor's elements in storage order (a chunk table narrows the
// binary search for the block that owns an element) and records the element's
// fused-matrix offset; for the operands the map is then inverted (fused
// offset -> blocked element, zero where no block covers the entry).  Every
// apply copies through these maps as a gather: one coalesced 4-byte map read
// and one coalesced write per destination element, the reads following the
// blocks' runs.
#include "heff_internal.h"

namespace graft {

static constexpr int kPermThreads = 256;

// Run once per plan: for every element
// of the blocked tensor its fused-matrix offset (kMapZero for a block without
// a fused rectangle, kMapGap between blocks).
__global__ void __launch_bounds__(kPermThreads)
    perm_map_kernel(const PermBlock* __restrict__ blocks, const int32_t* __restrict__ chunk_first,
                    int64_t total, uint32_t* __restrict__ map) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t c = e / kPermChunk;
    int lo = chunk_first[c], hi = chunk_first[c + 1];
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (blocks[mid].base <= e) lo = mid;
      else hi = mid - 1;
    }
    const PermBlock& b = blocks[lo];
    const int64_t local = e - b.base;   // blocks hold < 2^31 elements: 32-bit indexing below
    if (local < 0 || local >= b.size) {
      map[e] = kMapGap;
      continue;
    }
    if (b.other < 0) {
      map[e] = kMapZero;
      continue;
    }
    int64_t o = b.other;
    uint32_t loc = (uint32_t)local;
    for (int l = 0; l < b.rank; ++l) {
      const uint32_t d = (uint32_t)b.dims[l];
      const uint32_t q = loc / d;
      o += (int64_t)(loc - q * d) * b.ostr[l];
      loc = q;
    }
    map[e] = (uint32_t)o;
  }
}

// Inverse of a forward map (run once per plan): inv[map[e]] = e for every
// element with a fused offset; inv starts as kMapZero everywhere.
__global__ void __launch_bounds__(kPermThreads)
    invert_map_kernel(const uint32_t* __restrict__ map, int64_t total, uint32_t* __restrict__ inv) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const uint32_t o = map[e];
    if (o < kMapZero) inv[o] = (uint32_t)e;
  }
}

__global__ void __launch_bounds__(kPermThreads)
    fill_u32_kernel(uint32_t* __restrict__ x, int64_t n, uint32_t v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = v;
}

// The per-apply copy: dst[i] = src[map[i]] (zero for kMapZero, untouched for
// kMapGap).  Coalesced map read and write, no block search.
__global__ void __launch_bounds__(kPermThreads)
    map_gather_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                      const uint32_t* __restrict__ map, int64_t total) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint32_t o = map[i];
    if (o == kMapGap) continue;
    dst[i] = (o == kMapZero) ? double2{0.0, 0.0} : src[o];
  }
}

// Zero the hole rectangles of a fused operand matrix: one workgroup per
// rectangle, threads over its elements.
__global__ void __launch_bounds__(kPermThreads)
    zero_holes_kernel(const HoleRect* __restrict__ holes, int nholes, double2* __restrict__ fused) {
  for (int h = blockIdx.x; h < nholes; h += gridDim.x) {
    const HoleRect r = holes[h];
    const int n = r.rows * r.cols;
    for (int e = threadIdx.x; e < n; e += kPermThreads) {
      const int j = e / r.rows, i = e - j * r.rows;
      fused[r.base + (int64_t)i * r.rstr + (int64_t)j * r.cstr] = double2{0.0, 0.0};
    }
  }
}

__global__ void __launch_bounds__(kPermThreads)
    compose_map_kernel(const uint32_t* __restrict__ src, int64_t n, const uint32_t* __restrict__ fwd,
                       uint32_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t e = src[i];
    const uint32_t f = e < kMapZero ? fwd[e] : kMapZero;
    out[i] = f < kMapZero ? f : kMapZero;
  }
}

static int perm_grid(int64_t n) {
  int64_t g = (n + kPermThreads - 1) / kPermThreads;
  if (g > 16384) g = 16384;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_perm_map(const PermBlock* blocks, int nblocks, const int32_t* chunk_first,
                           int64_t total, uint32_t* map, hipStream_t stream) {
  if (total <= 0 || nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(perm_map_kernel, dim3(perm_grid(total)), dim3(kPermThreads), 0, stream, blocks,
                     chunk_first, total, map);
  return hipGetLastError();
}

hipError_t launch_invert_map(const uint32_t* map, int64_t total, uint32_t* inv, int64_t ninv,
                             hipStream_t stream) {
  if (ninv > 0)
    hipLaunchKernelGGL(fill_u32_kernel, dim3(perm_grid(ninv)), dim3(kPermThreads), 0, stream, inv,
                       ninv, kMapZero);
  if (total > 0)
    hipLaunchKernelGGL(invert_map_kernel, dim3(perm_grid(total)), dim3(kPermThreads), 0, stream,
                       map, total, inv);
  return hipGetLastError();
}

hipError_t launch_zero_holes(const HoleRect* holes, int nholes, void* fused, hipStream_t stream) {
  if (nholes <= 0) return hipSuccess;
  const int grid = nholes < 65536 ? nholes : 65536;
  hipLaunchKernelGGL(zero_holes_kernel, dim3(grid), dim3(kPermThreads), 0, stream, holes, nholes,
                     (double2*)fused);
  return hipGetLastError();
}

hipError_t launch_compose_map(const uint32_t* src, int64_t n, const uint32_t* fwd, uint32_t* out,
                              hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(compose_map_kernel, dim3(perm_grid(n)), dim3(kPermThreads), 0, stream, src, n,
                     fwd, out);
  return hipGetLastError();
}

hipError_t launch_map_gather(const void* src, void* dst, const uint32_t* map, int64_t total,
                             hipStream_t stream) {
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(map_gather_kernel, dim3(perm_grid(total)), dim3(kPermThreads), 0, stream,
                     (const double2*)src, (double2*)dst, map, total);
  return hipGetLastError();
}

}  // namespace graft
