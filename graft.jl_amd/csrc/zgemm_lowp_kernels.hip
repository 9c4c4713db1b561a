// Reduced-precision products for the same grouped strided complex GEMM:
//   GRAFT_PREC_FP32   operands rounded to FP32, v_mfma_f32_16x16x4f32,
//                     FP32 accumulation (four real products per complex one);
//   GRAFT_PREC_BF16X3 every FP32 operand split as hi + lo (two bf16 values),
//                     a*b ~ hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16,
//                     FP32 accumulation.
// Storage stays complex128 in HBM: loads convert when staging into LDS and
// the epilogue widens the FP32 accumulators.  Addressing (items, segments,
// offset decode, split-K, orientation, XCD order) is the one of the FP64
// kernel in zgemm_kernels.hip.  The f32-accumulating MFMAs leave
// D[4*kr + r][li] in element r of lane (16*kr + li) (measured).
#include "heff_internal.h"

namespace graft {

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

static constexpr uint32_t kBad = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t lp_decode(uint32_t idx, int rank, const int32_t* dims,
                                              const uint32_t* strides) {
  uint32_t off = 0;
  for (int l = 0; l < rank; ++l) {
    const uint32_t d = (uint32_t)dims[l];
    const uint32_t q = idx / d;
    off += (idx - q * d) * strides[l];
    idx = q;
  }
  return off;
}

// Split an FP32 value into bf16 hi + lo (round to nearest even both times).
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

__device__ __forceinline__ uint32_t pack2(__bf16 a, __bf16 b) {
  return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

// PL (BF16X3 only): A holds split-bf16 word pairs and B word quads at the
// element indices of the FP64 layout (written by the producing streaming
// step and by the B split), staged into LDS as they are: no conversion.
template <int BM, int BN, int WM, int WN, bool TR, int P, bool PL>
__global__ void __launch_bounds__(64 * (BM / WM) * (BN / WN))
    zgemm_lowp_kernel(const double2* __restrict__ Abase, const double2* __restrict__ Bbase,
                      double2* __restrict__ Cbase, const GemmItem* __restrict__ items, int nitems,
                      const GemmSeg* __restrict__ segs, const uint32_t* __restrict__ koff,
                      int total_tiles, int conjA, int conjB, int ksplit,
                      double2* __restrict__ partial) {
  // complex k per tile (32 for the bf16 form measured slower once tiles are
  // autotuned: its larger LDS stage halves the workgroups per CU)
  constexpr int BK = 16;
  constexpr bool BF = (P == GRAFT_PREC_BF16X3);
  static_assert(!PL || BF, "split-word operands are BF16x3 only");
  constexpr int NWN = BN / WN;
  constexpr int NT = 64 * (BM / WM) * NWN;
  constexpr int EA = BM * BK / NT;
  constexpr int EB = BN * BK / NT;
  constexpr int FM = WM / 16;
  constexpr int FN = WN / 16;
  static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0, "tile/thread mismatch");
  static_assert(EA >= 1 && EB >= 1, "tile too small for the thread count");
  // FP32: planes [k][re | im] for A, [k][re | im | -im] for B (+1 pad).
  constexpr int ASL = 2 * BM + 1, BSL = 3 * BN + 1;
  // BF16x3: rows of 2*BK bf16 (re, im interleaved) + 8 pad, i.e. 20 words.
  constexpr int RW = BK + 4;             // 32-bit words per row
  constexpr int LDS_F32 = BK * ASL + BK * BSL;                // floats
  constexpr int LDS_BF = (2 * BM + 4 * BN) * RW;               // 32-bit words
  __shared__ __attribute__((aligned(16))) uint32_t lds[BF ? LDS_BF : LDS_F32];
  float* As = reinterpret_cast<float*>(lds);
  float* Bs = As + BK * ASL;
  uint32_t* Ah = lds;                    // [BM][RW]
  uint32_t* Al = Ah + BM * RW;
  uint32_t* Brh = Al + BM * RW;          // [BN][RW] each
  uint32_t* Brl = Brh + BN * RW;
  uint32_t* Bih = Brl + BN * RW;
  uint32_t* Bil = Bih + BN * RW;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / NWN) * WM;
  const int wn0 = (wave % NWN) * WN;
  const int li = lane & 15;
  const int kr = lane >> 4;
  const int kcoff = kr >> 1;
  const int part = kr & 1;

  int tile, kc = 0;
  {
    const int nblk = total_tiles * ksplit;
    const int b = blockIdx.x;
    const int x = b & 7, i = b >> 3;
    const int q = nblk >> 3, r = nblk & 7;
    const int lin = x * q + (x < r ? x : r) + i;
    tile = lin / ksplit;
    kc = lin - tile * ksplit;
  }
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
  }
  const GemmItem* it = items + lo;
  const int M = it->M, N = it->N;
  const int mrank = it->mrank, nrank = it->nrank;
  const int local = tile - it->tile_begin;
  const int m0 = (local / it->tiles_n) * BM;
  const int n0 = (local % it->tiles_n) * BN;

  flt4 accR[FM][FN], accI[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      accR[a][b] = flt4{0.f, 0.f, 0.f, 0.f};
      accI[a][b] = flt4{0.f, 0.f, 0.f, 0.f};
    }
  const float sA = conjA ? -1.f : 1.f;
  const float sB = conjB ? -1.f : 1.f;

  const int seg_end = it->seg_begin + it->seg_count;
  for (int s = it->seg_begin; s < seg_end; ++s) {
    const GemmSeg* sg = segs + s;
    const double2* A = Abase + sg->a_off;
    const double2* B = Bbase + sg->b_off;
    const int K = sg->K;
    const uint32_t* kA = koff + sg->koff_idx;
    const uint32_t* kB = kA + K;
    const bool akf = sg->a_kfast != 0;
    const bool bkf = sg->b_kfast != 0;
    int kbeg = 0, kend = K;
    if (ksplit > 1) {
      const int chunk = ((K + BK - 1) / BK + ksplit - 1) / ksplit * BK;
      kbeg = kc * chunk;
      kend = min(K, kbeg + chunk);
    }
    uint32_t rowA[EA], colB[EB];
    if (akf) {
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int m = m0 + tid / BK + e * (NT / BK);
        rowA[e] = (m < M) ? lp_decode((uint32_t)m, mrank, it->mdims, sg->a_mstr) : kBad;
      }
    } else {
      const int m = m0 + tid % BM;
      rowA[0] = (m < M) ? lp_decode((uint32_t)m, mrank, it->mdims, sg->a_mstr) : kBad;
    }
    if (bkf) {
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int n = n0 + tid / BK + e * (NT / BK);
        colB[e] = (n < N) ? lp_decode((uint32_t)n, nrank, it->ndims, sg->b_nstr) : kBad;
      }
    } else {
      const int n = n0 + tid % BN;
      colB[0] = (n < N) ? lp_decode((uint32_t)n, nrank, it->ndims, sg->b_nstr) : kBad;
    }
    double2 ra[EA], rb[EB];
    uint2 wa[EA];
    uint4 wb[EB];
    const uint2* Aw = reinterpret_cast<const uint2*>(Abase) + sg->a_off;
    const uint4* Bw = reinterpret_cast<const uint4*>(Bbase) + sg->b_off;
    auto fetch = [&](int k0) {
      if constexpr (PL) {
      if (akf) {
        const int q = k0 + tid % BK;
        const uint32_t ko = (q < kend) ? kA[q] : 0u;
#pragma unroll
        for (int e = 0; e < EA; ++e)
          wa[e] = (q < kend && rowA[e] != kBad) ? Aw[(uint64_t)rowA[e] + ko] : uint2{0u, 0u};
      } else {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
          const int q = k0 + tid / BM + e * (NT / BM);
          wa[e] = (q < kend && rowA[0] != kBad) ? Aw[(uint64_t)rowA[0] + kA[q]] : uint2{0u, 0u};
        }
      }
      if (bkf) {
        const int q = k0 + tid % BK;
        const uint32_t ko = (q < kend) ? kB[q] : 0u;
#pragma unroll
        for (int e = 0; e < EB; ++e)
          wb[e] = (q < kend && colB[e] != kBad) ? Bw[(uint64_t)colB[e] + ko] : uint4{0u, 0u, 0u, 0u};
      } else {
#pragma unroll
        for (int e = 0; e < EB; ++e) {
          const int q = k0 + tid / BN + e * (NT / BN);
          wb[e] = (q < kend && colB[0] != kBad) ? Bw[(uint64_t)colB[0] + kB[q]] : uint4{0u, 0u, 0u, 0u};
        }
      }
            return;
      }
      if (akf) {
        const int q = k0 + tid % BK;
        const uint32_t ko = (q < kend) ? kA[q] : 0u;
#pragma unroll
        for (int e = 0; e < EA; ++e)
          ra[e] = (q < kend && rowA[e] != kBad) ? A[(uint64_t)rowA[e] + ko] : double2{0.0, 0.0};
      } else {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
          const int q = k0 + tid / BM + e * (NT / BM);
          ra[e] = (q < kend && rowA[0] != kBad) ? A[(uint64_t)rowA[0] + kA[q]] : double2{0.0, 0.0};
        }
      }
      if (bkf) {
        const int q = k0 + tid % BK;
        const uint32_t ko = (q < kend) ? kB[q] : 0u;
#pragma unroll
        for (int e = 0; e < EB; ++e)
          rb[e] = (q < kend && colB[e] != kBad) ? B[(uint64_t)colB[e] + ko] : double2{0.0, 0.0};
      } else {
#pragma unroll
        for (int e = 0; e < EB; ++e) {
          const int q = k0 + tid / BN + e * (NT / BN);
          rb[e] = (q < kend && colB[0] != kBad) ? B[(uint64_t)colB[0] + kB[q]] : double2{0.0, 0.0};
        }
      }
    };
    auto putA = [&](int p, int q, double2 v) {
      const float re = (float)v.x, im = sA * (float)v.y;
      if constexpr (BF) {
        __bf16 rh, rl, ih, il;
        split_bf16(re, rh, rl);
        split_bf16(im, ih, il);
        Ah[p * RW + q] = pack2(rh, ih);
        Al[p * RW + q] = pack2(rl, il);
      } else {
        As[q * ASL + p] = re;
        As[q * ASL + BM + p] = im;
      }
    };
    auto putB = [&](int q, int p, double2 v) {
      const float br = (float)v.x, bi = sB * (float)v.y;
      if constexpr (BF) {
        __bf16 rh, rl, ih, il, nih, nil;
        split_bf16(br, rh, rl);
        split_bf16(bi, ih, il);
        nih = (__bf16)(-(float)ih);
        nil = (__bf16)(-(float)il);
        Brh[p * RW + q] = pack2(rh, nih);   // (br, -bi) for the real part
        Brl[p * RW + q] = pack2(rl, nil);
        Bih[p * RW + q] = pack2(ih, rh);    // (bi, br) for the imaginary part
        Bil[p * RW + q] = pack2(il, rl);
      } else {
        Bs[q * BSL + p] = br;
        Bs[q * BSL + BN + p] = bi;
        Bs[q * BSL + 2 * BN + p] = -bi;
      }
    };
    // conjugation of a split-word A: the imaginary halves change sign
    const uint32_t cmaskA = conjA ? 0x80000000u : 0u;
    auto stage = [&]() {
      if constexpr (PL) {
      if (akf) {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
          const int p = tid / BK + e * (NT / BK), q = tid % BK;
          Ah[p * RW + q] = wa[e].x ^ cmaskA;
          Al[p * RW + q] = wa[e].y ^ cmaskA;
        }
      } else {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
          const int p = tid % BM, q = tid / BM + e * (NT / BM);
          Ah[p * RW + q] = wa[e].x ^ cmaskA;
          Al[p * RW + q] = wa[e].y ^ cmaskA;
        }
      }
      auto putw = [&](int q, int p, uint4 w) {
        Brh[p * RW + q] = w.x;
        Brl[p * RW + q] = w.y;
        Bih[p * RW + q] = w.z;
        Bil[p * RW + q] = w.w;
      };
      if (bkf) {
#pragma unroll
        for (int e = 0; e < EB; ++e) putw(tid % BK, tid / BK + e * (NT / BK), wb[e]);
      } else {
#pragma unroll
        for (int e = 0; e < EB; ++e) putw(tid / BN + e * (NT / BN), tid % BN, wb[e]);
      }
            return;
      }
      if (akf) {
#pragma unroll
        for (int e = 0; e < EA; ++e) putA(tid / BK + e * (NT / BK), tid % BK, ra[e]);
      } else {
#pragma unroll
        for (int e = 0; e < EA; ++e) putA(tid % BM, tid / BM + e * (NT / BM), ra[e]);
      }
      if (bkf) {
#pragma unroll
        for (int e = 0; e < EB; ++e) putB(tid % BK, tid / BK + e * (NT / BK), rb[e]);
      } else {
#pragma unroll
        for (int e = 0; e < EB; ++e) putB(tid / BN + e * (NT / BN), tid % BN, rb[e]);
      }
    };

    const int nkt = (kend + BK - 1) / BK;
    fetch(kbeg);
    for (int kt = kbeg / BK; kt < nkt; ++kt) {
      __syncthreads();
      stage();
      __syncthreads();
      if (kt + 1 < nkt) fetch((kt + 1) * BK);
      if constexpr (BF) {
#pragma unroll
       for (int ch = 0; ch < BK / 16; ++ch) {
        // One K' = 32 chunk (16 complex k): lane group kr holds complex k
        // 16 ch + 4kr .. 16 ch + 4kr + 3 (re, im pairs).
        const int kw = 16 * ch + 4 * kr;
        bf16x8 ah[FM], al[FM], brh[FN], brl[FN], bih[FN], bil[FN];
#pragma unroll
        for (int a = 0; a < FM; ++a) {
          const int row = wm0 + a * 16 + li;
          ah[a] = *reinterpret_cast<const bf16x8*>(Ah + row * RW + kw);
          al[a] = *reinterpret_cast<const bf16x8*>(Al + row * RW + kw);
        }
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          const int col = wn0 + b * 16 + li;
          brh[b] = *reinterpret_cast<const bf16x8*>(Brh + col * RW + kw);
          brl[b] = *reinterpret_cast<const bf16x8*>(Brl + col * RW + kw);
          bih[b] = *reinterpret_cast<const bf16x8*>(Bih + col * RW + kw);
          bil[b] = *reinterpret_cast<const bf16x8*>(Bil + col * RW + kw);
        }
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            if constexpr (TR) {
              accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(brh[b], ah[a], accR[a][b], 0, 0, 0);
              accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(brl[b], ah[a], accR[a][b], 0, 0, 0);
              accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(brh[b], al[a], accR[a][b], 0, 0, 0);
              accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bih[b], ah[a], accI[a][b], 0, 0, 0);
              accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bil[b], ah[a], accI[a][b], 0, 0, 0);
              accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bih[b], al[a], accI[a][b], 0, 0, 0);
            } else {
              accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], brh[b], accR[a][b], 0, 0, 0);
              accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], brl[b], accR[a][b], 0, 0, 0);
              accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], brh[b], accR[a][b], 0, 0, 0);
              accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bih[b], accI[a][b], 0, 0, 0);
              accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bil[b], accI[a][b], 0, 0, 0);
              accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bih[b], accI[a][b], 0, 0, 0);
            }
          }
       }
      } else {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
          const int kq = kk + kcoff;
          const float* ap = As + kq * ASL + part * BM + wm0 + li;
          const float* brp = Bs + kq * BSL + (part ? 2 * BN : 0) + wn0 + li;
          const float* bip = Bs + kq * BSL + (part ? 0 : BN) + wn0 + li;
          float av[FM], bre[FN], bim[FN];
#pragma unroll
          for (int a = 0; a < FM; ++a) av[a] = ap[a * 16];
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            bre[b] = brp[b * 16];
            bim[b] = bip[b * 16];
          }
#pragma unroll
          for (int a = 0; a < FM; ++a)
#pragma unroll
            for (int b = 0; b < FN; ++b) {
              if constexpr (TR) {
                accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bre[b], av[a], accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bim[b], av[a], accI[a][b], 0, 0, 0);
              } else {
                accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bre[b], accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bim[b], accI[a][b], 0, 0, 0);
              }
            }
        }
      }
    }
  }

  // Epilogue: f32-accumulating MFMAs leave D[4*kr + r][li]; D is the (m, n)
  // block, or (n, m) under TR.
  auto m_of = [&](int a, int r) { return m0 + wm0 + a * 16 + (TR ? li : 4 * kr + r); };
  auto n_of = [&](int b, int r) { return n0 + wn0 + b * 16 + (TR ? 4 * kr + r : li); };
  if (ksplit > 1) {
    double2* Pp = partial + (size_t)kc * N * M;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m_of(a, r);
        if (m >= M) continue;
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          const int n = n_of(b, r);
          if (n < N) Pp[(size_t)n * M + m] = double2{(double)accR[a][b][r], (double)accI[a][b][r]};
        }
      }
    return;
  }
  double2* C = Cbase + it->c_off;
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m_of(a, r);
      if (m >= M) continue;
      const uint32_t ro = lp_decode((uint32_t)m, mrank, it->mdims, it->c_mstr);
#pragma unroll
      for (int b = 0; b < FN; ++b) {
        const int n = n_of(b, r);
        if (n >= N) continue;
        const uint32_t co = lp_decode((uint32_t)n, nrank, it->ndims, it->c_nstr);
        C[(uint64_t)ro + co] = double2{(double)accR[a][b][r], (double)accI[a][b][r]};
      }
    }
}

template <int BM, int BN, int WM, int WN, int P>
static hipError_t launch_lp(const void* A, const void* B, void* C, const GemmItem* items,
                            int nitems, const GemmSeg* segs, const uint32_t* koff, int total_tiles,
                            int conjA, int conjB, int ksplit, void* partial, bool tr,
                            hipStream_t stream, bool planes) {
  constexpr int NT = 64 * (BM / WM) * (BN / WN);
  constexpr bool BF = (P == GRAFT_PREC_BF16X3);
  if (planes && (!BF || ksplit != 1)) return hipErrorInvalidValue;
#define GRAFT_LP(TRV, PLV)                                                                          \
  hipLaunchKernelGGL((zgemm_lowp_kernel<BM, BN, WM, WN, TRV, P, PLV>), dim3(total_tiles * ksplit),  \
                     dim3(NT), 0, stream, (const double2*)A, (const double2*)B, (double2*)C, items, \
                     nitems, segs, koff, total_tiles, conjA, conjB, ksplit, (double2*)partial)
  if constexpr (BF) {
    if (planes) {
      if (tr) GRAFT_LP(true, true); else GRAFT_LP(false, true);
      return hipGetLastError();
    }
  }
  if (tr) GRAFT_LP(true, false); else GRAFT_LP(false, false);
#undef GRAFT_LP
  return hipGetLastError();
}

// The reduced-precision kernels reuse the FP64 tile geometry of the chosen
// configuration (three-product configurations map to their four-product
// shape; the bf16 split already uses three products).
template <int P>
static hipError_t launch_lp_config(int config, const void* A, const void* B, void* C,
                                   const GemmItem* items, int nitems, const GemmSeg* segs,
                                   const uint32_t* koff, int total_tiles, int conjA, int conjB,
                                   int ksplit, void* partial, bool tr, hipStream_t stream,
                                   bool planes) {
  switch (config) {
    case 0: return launch_lp<128, 128, 64, 64, P>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, stream, planes);
    case 1: case 8: case 12: case 14: case 16: case 18: return launch_lp<128, 64, 64, 32, P>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, stream, planes);
    case 2: case 7: case 11: case 15: return launch_lp<64, 64, 32, 32, P>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, stream, planes);
    case 3: return launch_lp<256, 16, 64, 16, P>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, stream, planes);
    case 4: case 9: case 13: case 17: return launch_lp<64, 128, 32, 64, P>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, stream, planes);
    case 10: return launch_lp<64, 256, 64, 64, P>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, stream, planes);
    case 5: return launch_lp<32, 32, 16, 16, P>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, stream, planes);
    case 6: return launch_lp<16, 256, 16, 64, P>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, stream, planes);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_zgemm_lowp(int config, int precision, const void* A, const void* B, void* C,
                             const GemmItem* items, int nitems, const GemmSeg* segs,
                             const uint32_t* koff, int total_tiles, int conjA, int conjB,
                             int ksplit, void* partial, bool tr, hipStream_t stream, bool planes) {
  if (precision == GRAFT_PREC_FP32)
    return launch_lp_config<GRAFT_PREC_FP32>(config, A, B, C, items, nitems, segs, koff,
                                             total_tiles, conjA, conjB, ksplit, partial, tr, stream,
                                             planes);
  if (precision == GRAFT_PREC_BF16X3)
    return launch_lp_config<GRAFT_PREC_BF16X3>(config, A, B, C, items, nitems, segs, koff,
                                               total_tiles, conjA, conjB, ksplit, partial, tr,
                                               stream, planes);
  return hipErrorNotSupported;
}

}  // namespace graft
