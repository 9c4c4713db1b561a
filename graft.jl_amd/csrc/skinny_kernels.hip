// Streaming kernels for the HBM-bound TTNO-absorb steps (small K, one small
// open dimension): C[j, s] = sum_k big[j, k] * small[k, s] for every j of the
// large operand.  The small operand (K x S <= 16 x 64 complex, or <= 64 x 16
// in the accumulator variant) is staged once
// per workgroup in LDS and read as a broadcast; each thread owns one j,
// gathers its K inputs (consecutive threads read consecutive addresses when
// the large operand's first open leg has unit stride) and writes its S
// outputs (coalesced when the output's first leg is that same leg).  It
// consumes the same item/segment/koff description as the tiled MFMA GEMM.
#include "heff_internal.h"

namespace graft {

static constexpr int kSkinnyThreads = 256;

// Offset of flat index idx over legs (dims, strides), first leg fastest.
// Callers keep idx below the product of dims, so the last leg needs no
// division (a merged single-leg group needs none at all).
__device__ __forceinline__ uint32_t sk_decode(uint32_t idx, int rank, const int32_t* dims,
                                              const uint32_t* strides) {
  uint32_t off = 0;
  for (int l = 0; l + 1 < rank; ++l) {
    const uint32_t d = (uint32_t)dims[l];
    const uint32_t q = idx / d;
    off += (idx - q * d) * strides[l];
    idx = q;
  }
  if (rank > 0) off += idx * strides[rank - 1];
  return off;
}

// Split-bf16 operand words (GRAFT_PREC_BF16X3 consumers): an FP32 value x
// is hi + lo, hi = bf16(x), lo = bf16(x - hi), both rounded to nearest even;
// an element (re, im) becomes the words (re_hi | im_hi << 16) and
// (re_lo | im_lo << 16), the layout the bf16 GEMM stages.
__device__ __forceinline__ uint32_t sk_bf16_bits(float x) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)x);
}
__device__ __forceinline__ uint2 sk_split_words(double re, double im) {
  const float r = (float)re, i = (float)im;
  const __bf16 rh = (__bf16)r, ih = (__bf16)i;
  const float rl = r - (float)rh, il = i - (float)ih;
  return uint2{(uint32_t)__builtin_bit_cast(uint16_t, rh) | (sk_bf16_bits(i) << 16),
               sk_bf16_bits(rl) | (sk_bf16_bits(il) << 16)};
}
template <bool PL>
__device__ __forceinline__ void sk_store(double2* C, uint64_t e, double re, double im) {
  if constexpr (PL)
    reinterpret_cast<uint2*>(C)[e] = sk_split_words(re, im);
  else
    C[e] = double2{re, im};
}

// BIG_IS_A: A is the large operand (j = m, s = n); otherwise B is (j = n, s = m).
// PL: the result is written as split-bf16 word pairs (8 bytes per element at
// the element's index) for a BF16x3 consumer, instead of complex FP64.
template <bool BIG_IS_A, bool PL>
__global__ void __launch_bounds__(kSkinnyThreads)
    zskinny_f64_kernel(const double2* __restrict__ Abase, const double2* __restrict__ Bbase,
                       double2* __restrict__ Cbase, const GemmItem* __restrict__ item,
                       const GemmSeg* __restrict__ seg, const uint32_t* __restrict__ koff,
                       int conjA, int conjB) {
  __shared__ double2 sm[kSkinnyMaxK * kSkinnyMaxS];
  __shared__ uint32_t cso[kSkinnyMaxS];
  const int M = item->M, N = item->N, K = seg->K;
  const int S = BIG_IS_A ? N : M;
  const int64_t J = BIG_IS_A ? M : N;
  const double2* A = Abase + seg->a_off;
  const double2* B = Bbase + seg->b_off;
  const uint32_t* kA = koff + seg->koff_idx;
  const uint32_t* kB = kA + K;
  double2* C = Cbase + item->c_off;
  const double sgA = conjA ? -1.0 : 1.0, sgB = conjB ? -1.0 : 1.0;

  // Stage the small operand as sm[k * S + s], conjugation applied.
  for (int e = threadIdx.x; e < S * K; e += blockDim.x) {
    const int s = e % S, k = e / S;
    double2 v;
    if (BIG_IS_A) {
      v = B[sk_decode((uint32_t)s, item->nrank, item->ndims, seg->b_nstr) + kB[k]];
      v.y *= sgB;
    } else {
      v = A[sk_decode((uint32_t)s, item->mrank, item->mdims, seg->a_mstr) + kA[k]];
      v.y *= sgA;
    }
    sm[k * S + s] = v;
  }
  for (int s = threadIdx.x; s < S; s += blockDim.x)
    cso[s] = BIG_IS_A ? sk_decode((uint32_t)s, item->nrank, item->ndims, item->c_nstr)
                      : sk_decode((uint32_t)s, item->mrank, item->mdims, item->c_mstr);
  __syncthreads();

  const double sgBig = BIG_IS_A ? sgA : sgB;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < J;
       j += (int64_t)gridDim.x * blockDim.x) {
    uint32_t bo, co;
    const double2* big;
    const uint32_t* kb;
    if (BIG_IS_A) {
      bo = sk_decode((uint32_t)j, item->mrank, item->mdims, seg->a_mstr);
      co = sk_decode((uint32_t)j, item->mrank, item->mdims, item->c_mstr);
      big = A;
      kb = kA;
    } else {
      bo = sk_decode((uint32_t)j, item->nrank, item->ndims, seg->b_nstr);
      co = sk_decode((uint32_t)j, item->nrank, item->ndims, item->c_nstr);
      big = B;
      kb = kB;
    }
    double2 x[kSkinnyMaxK];
#pragma unroll
    for (int k = 0; k < kSkinnyMaxK; ++k) {
      if (k < K) {
        x[k] = big[(uint64_t)bo + kb[k]];
        x[k].y *= sgBig;
      }
    }
    for (int s = 0; s < S; ++s) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int k = 0; k < kSkinnyMaxK; ++k) {
        if (k < K) {
          const double2 w = sm[k * S + s];
          re += x[k].x * w.x - x[k].y * w.y;
          im += x[k].x * w.y + x[k].y * w.x;
        }
      }
      sk_store<PL>(C, (uint64_t)co + cso[s], re, im);
    }
  }
}

// Variant for longer K (<= kSkinnyAccMaxK) and a very small side
// (S <= kSkinnyAccMaxS): each thread keeps its S outputs in registers and
// streams its K inputs one at a time, so registers do not grow with K.
template <bool BIG_IS_A, bool PL>
__global__ void __launch_bounds__(kSkinnyThreads)
    zskinny_acc_f64_kernel(const double2* __restrict__ Abase, const double2* __restrict__ Bbase,
                           double2* __restrict__ Cbase, const GemmItem* __restrict__ item,
                           const GemmSeg* __restrict__ seg, const uint32_t* __restrict__ koff,
                           int conjA, int conjB) {
  __shared__ double2 sm[kSkinnyAccMaxK * kSkinnyAccMaxS];
  __shared__ uint32_t cso[kSkinnyAccMaxS];
  __shared__ uint32_t kbig[kSkinnyAccMaxK];
  const int M = item->M, N = item->N, K = seg->K;
  const int S = BIG_IS_A ? N : M;
  const int64_t J = BIG_IS_A ? M : N;
  const double2* A = Abase + seg->a_off;
  const double2* B = Bbase + seg->b_off;
  const uint32_t* kA = koff + seg->koff_idx;
  const uint32_t* kB = kA + K;
  double2* C = Cbase + item->c_off;
  const double sgA = conjA ? -1.0 : 1.0, sgB = conjB ? -1.0 : 1.0;

  for (int e = threadIdx.x; e < S * K; e += blockDim.x) {
    const int s = e % S, k = e / S;
    double2 v;
    if (BIG_IS_A) {
      v = B[sk_decode((uint32_t)s, item->nrank, item->ndims, seg->b_nstr) + kB[k]];
      v.y *= sgB;
    } else {
      v = A[sk_decode((uint32_t)s, item->mrank, item->mdims, seg->a_mstr) + kA[k]];
      v.y *= sgA;
    }
    sm[k * S + s] = v;
  }
  for (int s = threadIdx.x; s < S; s += blockDim.x)
    cso[s] = BIG_IS_A ? sk_decode((uint32_t)s, item->nrank, item->ndims, item->c_nstr)
                      : sk_decode((uint32_t)s, item->mrank, item->mdims, item->c_mstr);
  for (int k = threadIdx.x; k < K; k += blockDim.x) kbig[k] = BIG_IS_A ? kA[k] : kB[k];
  __syncthreads();

  const double sgBig = BIG_IS_A ? sgA : sgB;
  const double2* big = BIG_IS_A ? A : B;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < J;
       j += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t bo = BIG_IS_A ? sk_decode((uint32_t)j, item->mrank, item->mdims, seg->a_mstr)
                                 : sk_decode((uint32_t)j, item->nrank, item->ndims, seg->b_nstr);
    const uint32_t co = BIG_IS_A ? sk_decode((uint32_t)j, item->mrank, item->mdims, item->c_mstr)
                                 : sk_decode((uint32_t)j, item->nrank, item->ndims, item->c_nstr);
    double2 acc[kSkinnyAccMaxS];
#pragma unroll
    for (int s = 0; s < kSkinnyAccMaxS; ++s) acc[s] = double2{0.0, 0.0};
#pragma unroll 4
    for (int k = 0; k < K; ++k) {
      double2 x = big[(uint64_t)bo + kbig[k]];
      x.y *= sgBig;
#pragma unroll
      for (int s = 0; s < kSkinnyAccMaxS; ++s) {
        if (s < S) {
          const double2 w = sm[k * S + s];
          acc[s].x += x.x * w.x - x.y * w.y;
          acc[s].y += x.x * w.y + x.y * w.x;
        }
      }
    }
#pragma unroll
    for (int s = 0; s < kSkinnyAccMaxS; ++s)
      if (s < S) sk_store<PL>(C, (uint64_t)co + cso[s], acc[s].x, acc[s].y);
  }
}

// Per-block streaming for multi-block (sector) steps.  One wave per work
// entry, kBlockStreamLaneRows big-side elements j per lane (one after
// another), kBlockStreamS small-side outputs per pass held in registers.  Every address of the small operand, the koff
// tables and the item/segment records is the same across the wave (scalar
// loads); the big operand is read where its blocks lie, so no block is
// gathered into or scattered out of a fused matrix.  The output blocks of one
// entry share the big operand's open sector tuple: their rows are the same
// elements, and the reads after the first block hit the caches.
template <bool BIG_IS_A>
__global__ void __launch_bounds__(kSkinnyThreads)
    zstream_blocks_kernel(const double2* __restrict__ A, const double2* __restrict__ B,
                          double2* __restrict__ C, const GemmItem* __restrict__ items,
                          const GemmSeg* __restrict__ segs, const uint32_t* __restrict__ koff,
                          const int32_t* __restrict__ order, const StreamWork* __restrict__ work,
                          int nwork, int conjA, int conjB, const double* __restrict__ coef) {
  constexpr int kWaves = kSkinnyThreads / 64;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int w = (int)blockIdx.x * kWaves + wave;
  if (w >= nwork) return;
  const int lane = (int)(threadIdx.x & 63);
  const StreamWork wk = work[w];
  const double sgBig = (BIG_IS_A ? conjA : conjB) ? -1.0 : 1.0;
  const double sgSml = (BIG_IS_A ? conjB : conjA) ? -1.0 : 1.0;
  for (int t = 0; t < wk.count; ++t) {
    const GemmItem* it = items + order[wk.begin + t];
    const int S = BIG_IS_A ? it->N : it->M;
    const int J = BIG_IS_A ? it->M : it->N;
    const int brank = BIG_IS_A ? it->mrank : it->nrank;
    const int srank = BIG_IS_A ? it->nrank : it->mrank;
    const int32_t* bdims = BIG_IS_A ? it->mdims : it->ndims;
    const int32_t* sdims = BIG_IS_A ? it->ndims : it->mdims;
    const uint32_t* cbstr = BIG_IS_A ? it->c_mstr : it->c_nstr;
    const uint32_t* csstr = BIG_IS_A ? it->c_nstr : it->c_mstr;
    double2* Cb = C + it->c_off;
    for (int s0 = 0; s0 < S; s0 += kBlockStreamS) {
      const int sn = min(kBlockStreamS, S - s0);
      for (int r = 0; r < kBlockStreamLaneRows; ++r) {
        if (wk.j0 + 64 * r >= J) break;
        const int j = wk.j0 + 64 * r + lane;
        const bool live = j < J;
        double2 acc[kBlockStreamS];
#pragma unroll
        for (int s = 0; s < kBlockStreamS; ++s) acc[s] = double2{0.0, 0.0};
        for (int g = 0; g < it->seg_count; ++g) {
          const GemmSeg* sg = segs + it->seg_begin + g;
          const double cf = coef != nullptr ? coef[it->seg_begin + g] : 1.0;
          const double cfi = cf * sgBig;
          const int K = sg->K;
          const uint32_t* kA = koff + sg->koff_idx;
          const uint32_t* kbig = BIG_IS_A ? kA : kA + K;
          const uint32_t* ksml = BIG_IS_A ? kA + K : kA;
          const double2* big = BIG_IS_A ? A + sg->a_off : B + sg->b_off;
          const double2* sml = BIG_IS_A ? B + sg->b_off : A + sg->a_off;
          const uint32_t* bstr = BIG_IS_A ? sg->a_mstr : sg->b_nstr;
          const uint32_t* sstr = BIG_IS_A ? sg->b_nstr : sg->a_mstr;
          uint32_t so[kBlockStreamS];
#pragma unroll
          for (int s = 0; s < kBlockStreamS; ++s)
            so[s] = s < sn ? sk_decode((uint32_t)(s0 + s), srank, sdims, sstr) : 0u;
          if (live) {
            const uint32_t bo = sk_decode((uint32_t)j, brank, bdims, bstr);
#pragma unroll 4
            for (int k = 0; k < K; ++k) {
              double2 x = big[(uint64_t)bo + kbig[k]];
              x.x *= cf;
              x.y *= cfi;
              const uint32_t kk = ksml[k];
#pragma unroll
              for (int s = 0; s < kBlockStreamS; ++s) {
                if (s < sn) {
                  double2 v = sml[(uint64_t)so[s] + kk];
                  v.y *= sgSml;
                  acc[s].x += x.x * v.x - x.y * v.y;
                  acc[s].y += x.x * v.y + x.y * v.x;
                }
              }
            }
          }
        }
        if (live) {
          const uint32_t co = sk_decode((uint32_t)j, brank, bdims, cbstr);
#pragma unroll
          for (int s = 0; s < kBlockStreamS; ++s)
            if (s < sn) Cb[(uint64_t)co + sk_decode((uint32_t)(s0 + s), srank, sdims, csstr)] = acc[s];
        }
      }
    }
  }
}

// Lean per-block streaming: each lane owns kLeanRows big-side rows (j0 +
// lane + 64 r) of one output block, every small-side output of those rows in
// registers; the rows' loads are independent, so a wave keeps kLeanRows times
// as many loads in flight (the step is latency-bound).  The workgroup stages
// the whole small operand (conjugation applied) in LDS once; its waves then
// walk work entries with a grid stride.  An entry is 64 * kLeanRows rows of
// one output block (or of every block of one big-side tuple, on very large
// steps).  Row digits come from a multiply-high by the per-leg reciprocal
// the host computed.  Item and segment records are fixed-width and read as
// whole 16-byte vectors before any use, so the wave waits once per record.
template <bool BIG_IS_A, int SM, bool SMEM>
__global__ void __launch_bounds__(kSkinnyThreads)
    zstream_lean_kernel(const double2* __restrict__ A, const double2* __restrict__ B,
                        double2* __restrict__ C, const LeanItem* __restrict__ items,
                        const LeanSeg* __restrict__ segs, const LeanWork* __restrict__ work,
                        int nwork, int nsmall, int conjA, int conjB) {
  extern __shared__ double2 sm[];
  constexpr int kWaves = kSkinnyThreads / 64;
  constexpr int RR = kLeanRows;
  static_assert(kLeanRank == 4 && kLeanInlineK == 8 && SM % 4 == 0, "record layout");
  const double2* __restrict__ small = BIG_IS_A ? B : A;
  const double2* __restrict__ big = BIG_IS_A ? A : B;
  const double sgSml = (BIG_IS_A ? conjB : conjA) ? -1.0 : 1.0;
  const double sgBig = (BIG_IS_A ? conjA : conjB) ? -1.0 : 1.0;
  if constexpr (SMEM) {
    for (int i = threadIdx.x; i < nsmall; i += kSkinnyThreads) {
      double2 v = small[i];
      v.y *= sgSml;
      sm[i] = v;
    }
    __syncthreads();
  }
  const int lane = (int)(threadIdx.x & 63);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int nw = (int)gridDim.x * kWaves;
  for (int w = (int)blockIdx.x * kWaves + wave; w < nwork; w += nw) {
    const LeanWork wk = work[w];
    const LeanItem* it0 = items + wk.begin;
    const int J = it0->J, rank = it0->rank;
    const uint4 dims = *reinterpret_cast<const uint4*>(it0->dims);
    const uint4 magic = *reinterpret_cast<const uint4*>(it0->magic);
    if (wk.j0 + lane >= J) continue;
    // digits of each row over the big-side legs (legs past rank get 0); a
    // row past J takes row j0 + lane's digits and stores nothing
    bool live[RR];
    uint32_t d0[RR], d1[RR], d2[RR], d3[RR];
#pragma unroll
    for (int r = 0; r < RR; ++r) {
      const int jr = wk.j0 + lane + 64 * r;
      live[r] = jr < J;
      d0[r] = (uint32_t)(live[r] ? jr : wk.j0 + lane);
      d1[r] = d2[r] = d3[r] = 0u;
      if (rank > 1) {
        const uint32_t q0 = __umulhi(d0[r], magic.x);
        d1[r] = q0;
        d0[r] -= q0 * dims.x;
        if (rank > 2) {
          const uint32_t q1 = __umulhi(d1[r], magic.y);
          d2[r] = q1;
          d1[r] -= q1 * dims.y;
          if (rank > 3) {
            const uint32_t q2 = __umulhi(d2[r], magic.z);
            d3[r] = q2;
            d2[r] -= q2 * dims.z;
          }
        }
      }
    }
    for (int t = 0; t < wk.count; ++t) {
      const LeanItem* it = it0 + t;
      const int S = it->S;
      const uint4 cstr = *reinterpret_cast<const uint4*>(it->cstr);
      uint4 cso[SM / 4];
#pragma unroll
      for (int q = 0; q < SM / 4; ++q) cso[q] = reinterpret_cast<const uint4*>(it->cso)[q];
      double2 acc[RR][SM];
#pragma unroll
      for (int r = 0; r < RR; ++r)
#pragma unroll
        for (int s = 0; s < SM; ++s) acc[r][s] = double2{0.0, 0.0};
      const int gend = it->seg_begin + it->seg_count;
      for (int g = it->seg_begin; g < gend; ++g) {
        const LeanSeg* sg = segs + g;
        const uint4 bstr = *reinterpret_cast<const uint4*>(sg->bstr);
        const uint4 kb0 = reinterpret_cast<const uint4*>(sg->kb)[0];
        const uint4 kb1 = reinterpret_cast<const uint4*>(sg->kb)[1];
        const uint4 ks0 = reinterpret_cast<const uint4*>(sg->ks)[0];
        const uint4 ks1 = reinterpret_cast<const uint4*>(sg->ks)[1];
        uint4 sso[SM / 4];
#pragma unroll
        for (int q = 0; q < SM / 4; ++q) sso[q] = reinterpret_cast<const uint4*>(sg->sso)[q];
        const int K = sg->K;
        const double cf = sg->coef;
        const double cfi = cf * sgBig;
        const uint32_t kbv[8] = {kb0.x, kb0.y, kb0.z, kb0.w, kb1.x, kb1.y, kb1.z, kb1.w};
        const uint32_t ksv[8] = {ks0.x, ks0.y, ks0.z, ks0.w, ks1.x, ks1.y, ks1.z, ks1.w};
        const double2* bp[RR];
#pragma unroll
        for (int r = 0; r < RR; ++r)
          bp[r] = big + sg->big_off +
                  (d0[r] * bstr.x + d1[r] * bstr.y + d2[r] * bstr.z + d3[r] * bstr.w);
#pragma unroll
        for (int k = 0; k < kLeanInlineK; ++k) {
          if (k < K) {
            double2 x[RR];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
              x[r] = bp[r][kbv[k]];
              x[r].x *= cf;
              x[r].y *= cfi;
            }
#pragma unroll
            for (int s = 0; s < SM; ++s) {
              if (s < S) {
                const uint32_t so = (&sso[s / 4].x)[s % 4];
                double2 v;
                if constexpr (SMEM) {
                  v = sm[ksv[k] + so];
                } else {
                  // uniform address: one scalar load for the wave, cached
                  v = small[(uint64_t)ksv[k] + so];
                  v.y *= sgSml;
                }
#pragma unroll
                for (int r = 0; r < RR; ++r) {
                  acc[r][s].x += x[r].x * v.x - x[r].y * v.y;
                  acc[r][s].y += x[r].x * v.y + x[r].y * v.x;
                }
              }
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < RR; ++r) {
        if (!live[r]) continue;
        double2* cp =
            C + it->c_off + (d0[r] * cstr.x + d1[r] * cstr.y + d2[r] * cstr.z + d3[r] * cstr.w);
#pragma unroll
        for (int s = 0; s < SM; ++s)
          if (s < S) cp[(&cso[s / 4].x)[s % 4]] = acc[r][s];
      }
    }
  }
}

hipError_t launch_zstream_lean(bool big_is_a, int smax, const void* A, const void* B, void* C,
                               const LeanItem* items, const LeanSeg* segs, const LeanWork* work,
                               int nwork, int nsmall, int conjA, int conjB, hipStream_t stream) {
  if (nwork <= 0) return hipSuccess;
  if (nsmall < 0 || smax < 1 || smax > kLeanMaxS) return hipErrorInvalidValue;
  // a small operand that fits in LDS is staged there; a larger one is read
  // from global memory at wave-uniform addresses (cached scalar loads)
  const bool smem = nsmall <= kLeanSmallMax;
  constexpr int kWaves = kSkinnyThreads / 64;
  // enough workgroups to fill the chip several times over; each stages the
  // small operand once, so fewer and longer-lived when it is large
  const int cap = smem && nsmall > 1024 ? 2048 : 8192;
  const int grid = std::min((nwork + kWaves - 1) / kWaves, cap);
  const size_t lds = smem ? (size_t)std::max(nsmall, 1) * sizeof(double2) : 0;
  const double2* a = (const double2*)A;
  const double2* b = (const double2*)B;
  double2* c = (double2*)C;
#define GRAFT_LEAN(BIGA, SMV)                                                                      \
  do {                                                                                             \
    if (smem)                                                                                      \
      hipLaunchKernelGGL((zstream_lean_kernel<BIGA, SMV, true>), dim3(grid), dim3(kSkinnyThreads), \
                         lds, stream, a, b, c, items, segs, work, nwork, nsmall, conjA, conjB);    \
    else                                                                                           \
      hipLaunchKernelGGL((zstream_lean_kernel<BIGA, SMV, false>), dim3(grid), dim3(kSkinnyThreads),\
                         lds, stream, a, b, c, items, segs, work, nwork, nsmall, conjA, conjB);    \
  } while (0)
  if (big_is_a) {
    if (smax <= 4) GRAFT_LEAN(true, 4);
    else if (smax <= 8) GRAFT_LEAN(true, 8);
    else GRAFT_LEAN(true, 16);
  } else {
    if (smax <= 4) GRAFT_LEAN(false, 4);
    else if (smax <= 8) GRAFT_LEAN(false, 8);
    else GRAFT_LEAN(false, 16);
  }
#undef GRAFT_LEAN
  return hipGetLastError();
}

hipError_t launch_zstream_blocks(bool big_is_a, const void* A, const void* B, void* C,
                                 const GemmItem* items, const GemmSeg* segs, const uint32_t* koff,
                                 const int32_t* order, const StreamWork* work, int nwork, int conjA,
                                 int conjB, hipStream_t stream, const double* coef) {
  if (nwork <= 0) return hipSuccess;
  constexpr int kWaves = kSkinnyThreads / 64;
  const int grid = (nwork + kWaves - 1) / kWaves;
  if (big_is_a)
    hipLaunchKernelGGL(zstream_blocks_kernel<true>, dim3(grid), dim3(kSkinnyThreads), 0, stream,
                       (const double2*)A, (const double2*)B, (double2*)C, items, segs, koff, order,
                       work, nwork, conjA, conjB, coef);
  else
    hipLaunchKernelGGL(zstream_blocks_kernel<false>, dim3(grid), dim3(kSkinnyThreads), 0, stream,
                       (const double2*)A, (const double2*)B, (double2*)C, items, segs, koff, order,
                       work, nwork, conjA, conjB, coef);
  return hipGetLastError();
}

int skinny_grid(int64_t J) {
  int64_t g = (J + kSkinnyThreads - 1) / kSkinnyThreads;
  if (g > 16384) g = 16384;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_zskinny(bool big_is_a, bool acc, const void* A, const void* B, void* C,
                          const GemmItem* item, const GemmSeg* seg, const uint32_t* koff, int grid,
                          int conjA, int conjB, hipStream_t stream, bool planes) {
  const double2* a = (const double2*)A;
  const double2* b = (const double2*)B;
  double2* c = (double2*)C;
#define GRAFT_SK(KER, BIGA, PLV)                                                                    \
  hipLaunchKernelGGL((KER<BIGA, PLV>), dim3(grid), dim3(kSkinnyThreads), 0, stream, a, b, c, item, \
                     seg, koff, conjA, conjB)
#define GRAFT_SK2(KER, BIGA)                                                                        \
  if (planes) GRAFT_SK(KER, BIGA, true); else GRAFT_SK(KER, BIGA, false)
  if (acc) {
    if (big_is_a) { GRAFT_SK2(zskinny_acc_f64_kernel, true); }
    else { GRAFT_SK2(zskinny_acc_f64_kernel, false); }
  } else {
    if (big_is_a) { GRAFT_SK2(zskinny_f64_kernel, true); }
    else { GRAFT_SK2(zskinny_f64_kernel, false); }
  }
#undef GRAFT_SK2
#undef GRAFT_SK
  return hipGetLastError();
}

// Split-bf16 words of a complex FP64 operand for the bf16 GEMM's B side,
// conjugation applied: per element (br_h | -bi_h), (br_l | -bi_l),
// (bi_h | br_h), (bi_l | br_l) (the real-part and imaginary-part columns
// of the four-product form).
__global__ void __launch_bounds__(kSkinnyThreads)
    split_bplanes_kernel(const double2* __restrict__ B, int64_t n, uint4* __restrict__ out, int conjB) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2 v = B[i];
    const uint2 w = sk_split_words(v.x, conjB ? -v.y : v.y);
    const uint32_t rh = w.x & 0xFFFFu, ih = w.x >> 16, rl = w.y & 0xFFFFu, il = w.y >> 16;
    out[i] = uint4{rh | ((ih ^ 0x8000u) << 16), rl | ((il ^ 0x8000u) << 16), ih | (rh << 16),
                   il | (rl << 16)};
  }
}

hipError_t launch_split_bplanes(const void* B, int64_t n, void* out, int conjB, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  int64_t g = (n + kSkinnyThreads - 1) / kSkinnyThreads;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL(split_bplanes_kernel, dim3((int)g), dim3(kSkinnyThreads), 0, stream,
                     (const double2*)B, n, (uint4*)out, conjB);
  return hipGetLastError();
}

}  // namespace graft
