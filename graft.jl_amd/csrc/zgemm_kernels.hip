// Grouped, strided, permute-fused complex GEMM on CDNA4 (gfx950) MFMA.
//
// One launch executes one pairwise contraction step of an effective-
// Hamiltonian or environment plan: for every output block (GemmItem) and
// every (A block, B block) contribution to it (GemmSeg),
//     C[m, n] = sum_seg sum_k op(A)[m, k] * op(B)[k, n].
// Legs are never permuted in memory.  Each operand element address is the
// sum of a row part and a column part, because the open and the contracted
// legs of a tensor are disjoint: offset(m, k) = rowoff(m) + koff(k).  Row
// offsets are decoded once per tile from the merged leg dims and strides;
// contracted offsets come from a per-step table.  That folds the permutes
// TensorOperations performs around its zgemm into the tile loads/stores.
//
// Complex arithmetic on real MFMA: the 16x16x4 f64 instruction consumes four
// real k values per lane group.  Two complex k values are laid out as
// (k0.re, k0.im, k1.re, k1.im), so with A in that interleaved form,
//     Cr = A_ri . [ Br; -Bi ],   Ci = A_ri . [ Bi; Br ],
// i.e. two MFMAs per 16x16 complex block and complex k pair (4 real
// products per complex product).  The B tile is staged in LDS as three
// planes (Br, Bi, -Bi) so that every lane reads its operand directly.
// Conjugation flags are folded into the LDS stores.
#include "heff_internal.h"

namespace graft {

const TileConfig kTileConfigs[kNumTileConfigs] = {
    {128, 128, 16, 64, 64, 0},  // 0: large m and n
    {128, 64, 16, 64, 32, 0},   // 1
    {64, 64, 16, 32, 32, 0},    // 2
    {256, 16, 16, 64, 16, 0},   // 3: skinny n (TTNO absorb)
    {64, 128, 16, 32, 64, 0},   // 4
    {32, 32, 16, 16, 16, 0},    // 5: small blocks
    {16, 256, 16, 16, 64, 0},   // 6: skinny m (environment closure)
    {64, 64, 16, 32, 32, 1},    // 7: three-product form
    {128, 64, 16, 64, 32, 1},   // 8: three-product form
    {64, 128, 16, 32, 64, 1},   // 9: three-product form
    {64, 256, 8, 64, 64, 0},    // 10: n up to 256 in one tile (each A panel read once)
    {64, 64, 16, 32, 32, 1},    // 11: three-product form, double-buffered LDS
    {128, 64, 16, 64, 32, 1},   // 12: three-product form, double-buffered LDS
    {64, 128, 16, 32, 64, 1},   // 13: three-product form, double-buffered LDS
    {128, 64, 16, 32, 32, 1},   // 14: three-product form, double-buffered LDS, 8 waves
    {64, 64, 16, 32, 32, 1},    // 15-18: 11-14 with the pipelined kernel on whole-tile steps
    {128, 64, 16, 64, 32, 1},   // 16
    {64, 128, 16, 32, 64, 1},   // 17
    {128, 64, 16, 32, 32, 1},   // 18
};

typedef double dbl4 __attribute__((ext_vector_type(4)));

static constexpr uint32_t kBadRow = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t decode_offset(uint32_t idx, int rank, const int32_t* dims,
                                                  const uint32_t* strides) {
  uint32_t off = 0;
  for (int l = 0; l < rank; ++l) {
    uint32_t d = (uint32_t)dims[l];
    uint32_t q = idx / d;
    off += (idx - q * d) * strides[l];
    idx = q;
  }
  return off;
}

// TR (transposed output orientation): issue the MFMAs as B^T A^T so that the
// 16 lanes of an accumulator column walk the m index.  Chosen when the
// output's unit-stride leg is an m leg, which makes the stores coalesced.
//
// M3 (three-product form): T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi),
// Cr = T1 - T2, Ci = T3 - T1 - T2 — three real MFMAs per four complex k and
// 16x16 block instead of four (Karatsuba / Gauss).  In FP64 its rounding
// error stays O(eps (|Ar| + |Ai|)(|Br| + |Bi|)) per term.
//
// MAP (fused sector steps): the item/segment offsets address per-charge fused
// matrices.  An operand whose map is given is never formed: maps.a / maps.b
// give the blocked element of every fused A / B entry (kMapZero: no block,
// the entry is zero) and maps.c the blocked element each fused C entry is
// stored to (kMapZero: not stored); an operand without a map is read from
// its gathered fused matrix as usual.
template <int BM, int BN, int BK, int WM, int WN, bool TR, bool M3, bool MAP>
__global__ void __launch_bounds__(64 * (BM / WM) * (BN / WN))
    zgemm_f64_kernel(const double2* __restrict__ Abase, const double2* __restrict__ Bbase,
                     double2* __restrict__ Cbase, const GemmItem* __restrict__ items, int nitems,
                     const GemmSeg* __restrict__ segs, const uint32_t* __restrict__ koff,
                     int total_tiles, int conjA, int conjB, int ksplit,
                     double2* __restrict__ partial, FusedMaps maps) {
  constexpr int NWN = BN / WN;
  constexpr int NT = 64 * (BM / WM) * NWN;
  constexpr int EA = BM * BK / NT;
  constexpr int EB = BN * BK / NT;
  constexpr int FM = WM / 16;
  constexpr int FN = WN / 16;
  // LDS slab (doubles) per k.  Four-product form: A (re, im), B (re, im, -im);
  // three-product form: A (re, im, re+im), B (re, im, re+im).  +1 pad.
  constexpr int ASL = (M3 ? 3 : 2) * BM + 1;
  constexpr int BSL = 3 * BN + 1;
  constexpr int NA = M3 ? 3 : 2;   // accumulators per 16x16 block
  static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0, "tile/thread mismatch");
  static_assert(EA >= 1 && EB >= 1, "tile too small for the thread count");

  __shared__ double As[BK * ASL];
  __shared__ double Bs[BK * BSL];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / NWN) * WM;
  const int wn0 = (wave % NWN) * WN;
  const int li = lane & 15;
  const int kr = lane >> 4;
  const int kcoff = kr >> 1;
  const int part = kr & 1;

  // XCD-aware tile order: hardware dispatches block b to XCD b % 8, so give
  // each XCD a contiguous range of logical tiles (tiles sharing an A panel
  // then share that XCD's L2).
  // With split-K (ksplit > 1, single-block items only) each logical block
  // handles one k-chunk of one tile and writes a partial tile.
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

  dbl4 acc[NA][FM][FN];
#pragma unroll
  for (int t = 0; t < NA; ++t)
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) acc[t][a][b] = dbl4{0.0, 0.0, 0.0, 0.0};

  const double sA = conjA ? -1.0 : 1.0;
  const double sB = conjB ? -1.0 : 1.0;

  const int seg_end = it->seg_begin + it->seg_count;
  for (int s = it->seg_begin; s < seg_end; ++s) {
    const GemmSeg* sg = segs + s;
    const bool mapa = MAP && maps.a != nullptr, mapb = MAP && maps.b != nullptr;
    const double2* A = Abase + (mapa ? 0 : sg->a_off);
    const double2* B = Bbase + (mapb ? 0 : sg->b_off);
    const uint32_t* mA = mapa ? maps.a + sg->a_off : nullptr;
    const uint32_t* mB = mapb ? maps.b + sg->b_off : nullptr;
    auto ldA = [&](uint64_t i) -> double2 {
      if constexpr (MAP) {
        if (mapa) {
          const uint32_t e = mA[i];
          return e < kMapZero ? A[e] : double2{0.0, 0.0};
        }
      }
      return A[i];
    };
    auto ldB = [&](uint64_t i) -> double2 {
      if constexpr (MAP) {
        if (mapb) {
          const uint32_t e = mB[i];
          return e < kMapZero ? B[e] : double2{0.0, 0.0};
        }
      }
      return B[i];
    };
    // recoupling factor of this block pair (coupled steps); 1 otherwise
    const double ca = (MAP && maps.coef != nullptr) ? maps.coef[s] : 1.0;
    const double caI = ca * sA;
    const int K = sg->K;
    const uint32_t* kA = koff + sg->koff_idx;
    const uint32_t* kB = kA + K;
    const bool akf = sg->a_kfast != 0;
    const bool bkf = sg->b_kfast != 0;
    // k range of this block: all of K, or one BK-aligned chunk under split-K.
    int kbeg = 0, kend = K;
    if (ksplit > 1) {
      const int chunk = ((K + BK - 1) / BK + ksplit - 1) / ksplit * BK;
      kbeg = kc * chunk;
      kend = min(K, kbeg + chunk);
    }

    // Row offsets of this thread's A elements and column offsets of its B elements.
    uint32_t rowA[EA], colB[EB];
    if (akf) {
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int p = tid / BK + e * (NT / BK);
        const int m = m0 + p;
        rowA[e] = (m < M) ? decode_offset((uint32_t)m, mrank, it->mdims, sg->a_mstr) : kBadRow;
      }
    } else {
      const int m = m0 + tid % BM;
      rowA[0] = (m < M) ? decode_offset((uint32_t)m, mrank, it->mdims, sg->a_mstr) : kBadRow;
    }
    if (bkf) {
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int p = tid / BK + e * (NT / BK);
        const int n = n0 + p;
        colB[e] = (n < N) ? decode_offset((uint32_t)n, nrank, it->ndims, sg->b_nstr) : kBadRow;
      }
    } else {
      const int n = n0 + tid % BN;
      colB[0] = (n < N) ? decode_offset((uint32_t)n, nrank, it->ndims, sg->b_nstr) : kBadRow;
    }

    double2 ra[EA], rb[EB];
    auto fetch = [&](int k0) {
      if (akf) {
        const int q = k0 + tid % BK;
        const uint32_t ko = (q < kend) ? kA[q] : 0u;
#pragma unroll
        for (int e = 0; e < EA; ++e)
          ra[e] = (q < kend && rowA[e] != kBadRow) ? ldA((uint64_t)rowA[e] + ko) : double2{0.0, 0.0};
      } else {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
          const int q = k0 + tid / BM + e * (NT / BM);
          ra[e] = (q < kend && rowA[0] != kBadRow) ? ldA((uint64_t)rowA[0] + kA[q]) : double2{0.0, 0.0};
        }
      }
      if (bkf) {
        const int q = k0 + tid % BK;
        const uint32_t ko = (q < kend) ? kB[q] : 0u;
#pragma unroll
        for (int e = 0; e < EB; ++e)
          rb[e] = (q < kend && colB[e] != kBadRow) ? ldB((uint64_t)colB[e] + ko) : double2{0.0, 0.0};
      } else {
#pragma unroll
        for (int e = 0; e < EB; ++e) {
          const int q = k0 + tid / BN + e * (NT / BN);
          rb[e] = (q < kend && colB[0] != kBadRow) ? ldB((uint64_t)colB[0] + kB[q]) : double2{0.0, 0.0};
        }
      }
    };
    auto stage = [&]() {
      if (akf) {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
          const int q = tid % BK, p = tid / BK + e * (NT / BK);
          const double re = MAP ? ca * ra[e].x : ra[e].x;
          const double im = MAP ? caI * ra[e].y : sA * ra[e].y;
          As[q * ASL + p] = re;
          As[q * ASL + BM + p] = im;
          if constexpr (M3) As[q * ASL + 2 * BM + p] = re + im;
        }
      } else {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
          const int p = tid % BM, q = tid / BM + e * (NT / BM);
          const double re = MAP ? ca * ra[e].x : ra[e].x;
          const double im = MAP ? caI * ra[e].y : sA * ra[e].y;
          As[q * ASL + p] = re;
          As[q * ASL + BM + p] = im;
          if constexpr (M3) As[q * ASL + 2 * BM + p] = re + im;
        }
      }
      if (bkf) {
#pragma unroll
        for (int e = 0; e < EB; ++e) {
          const int q = tid % BK, p = tid / BK + e * (NT / BK);
          const double im = sB * rb[e].y;
          Bs[q * BSL + p] = rb[e].x;
          Bs[q * BSL + BN + p] = im;
          Bs[q * BSL + 2 * BN + p] = M3 ? rb[e].x + im : -im;
        }
      } else {
#pragma unroll
        for (int e = 0; e < EB; ++e) {
          const int p = tid % BN, q = tid / BN + e * (NT / BN);
          const double im = sB * rb[e].y;
          Bs[q * BSL + p] = rb[e].x;
          Bs[q * BSL + BN + p] = im;
          Bs[q * BSL + 2 * BN + p] = M3 ? rb[e].x + im : -im;
        }
      }
    };

    const int nkt = (kend + BK - 1) / BK;
    fetch(kbeg);
    for (int kt = kbeg / BK; kt < nkt; ++kt) {
      __syncthreads();
      stage();
      __syncthreads();
      if (kt + 1 < nkt) fetch((kt + 1) * BK);
      if constexpr (M3) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
          const int kq = kk + kr;   // lane group kr supplies complex k = kk + kr
          double av[3][FM], bv[3][FN];
#pragma unroll
          for (int t = 0; t < 3; ++t) {
#pragma unroll
            for (int a = 0; a < FM; ++a) av[t][a] = As[kq * ASL + t * BM + wm0 + a * 16 + li];
#pragma unroll
            for (int b = 0; b < FN; ++b) bv[t][b] = Bs[kq * BSL + t * BN + wn0 + b * 16 + li];
          }
#pragma unroll
          for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int a = 0; a < FM; ++a)
#pragma unroll
              for (int b = 0; b < FN; ++b) {
                if constexpr (TR)
                  acc[t][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[t][b], av[t][a], acc[t][a][b], 0, 0, 0);
                else
                  acc[t][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t][a], bv[t][b], acc[t][a][b], 0, 0, 0);
              }
        }
      } else {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
          const int kq = kk + kcoff;
          const double* ap = As + kq * ASL + part * BM + wm0 + li;
          const double* brp = Bs + kq * BSL + (part ? 2 * BN : 0) + wn0 + li;
          const double* bip = Bs + kq * BSL + (part ? 0 : BN) + wn0 + li;
          double av[FM], bre[FN], bim[FN];
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
                acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bre[b], av[a], acc[0][a][b], 0, 0, 0);
                acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bim[b], av[a], acc[1][a][b], 0, 0, 0);
              } else {
                acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bre[b], acc[0][a][b], 0, 0, 0);
                acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bim[b], acc[1][a][b], 0, 0, 0);
              }
            }
        }
      }
    }
  }

  // Epilogue: v_mfma_f64_16x16x4f64 leaves D[kr + 4*r][li] in accumulator
  // element r of lane (16*kr + li); D is the (m, n) block, or (n, m) under TR.
  auto m_of = [&](int a, int r) { return m0 + wm0 + a * 16 + (TR ? li : kr + 4 * r); };
  auto cval = [&](int a, int b, int r) -> double2 {
    if constexpr (M3) {
      const double t1 = acc[0][a][b][r], t2 = acc[1][a][b][r], t3 = acc[2][a][b][r];
      return double2{t1 - t2, t3 - t1 - t2};
    } else {
      return double2{acc[0][a][b][r], acc[1][a][b][r]};
    }
  };
  auto n_of = [&](int b, int r) { return n0 + wn0 + b * 16 + (TR ? kr + 4 * r : li); };
  if (ksplit > 1) {
    // partial tile of k-chunk kc, element (m, n) at partial[(kc * N + n) * M + m]
    double2* P = partial + (size_t)kc * N * M;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m_of(a, r);
        if (m >= M) continue;
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          const int n = n_of(b, r);
          if (n < N) P[(size_t)n * M + m] = cval(a, b, r);
        }
      }
    return;
  }
  const bool mapc = MAP && maps.c != nullptr;
  double2* C = Cbase + (mapc ? 0 : it->c_off);
  const uint32_t* mC = mapc ? maps.c + it->c_off : nullptr;
  auto stC = [&](uint64_t i, double2 v) {
    if constexpr (MAP) {
      if (mapc) {
        const uint32_t e = mC[i];
        if (e < kMapZero) C[e] = v;
        return;
      }
    }
    C[i] = v;
  };
  if constexpr (TR) {
    uint32_t rowm[FM];
#pragma unroll
    for (int a = 0; a < FM; ++a) {
      const int m = m_of(a, 0);
      rowm[a] = (m < M) ? decode_offset((uint32_t)m, mrank, it->mdims, it->c_mstr) : kBadRow;
    }
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n_of(b, r);
        if (n >= N) continue;
        const uint32_t co = decode_offset((uint32_t)n, nrank, it->ndims, it->c_nstr);
#pragma unroll
        for (int a = 0; a < FM; ++a) {
          if (rowm[a] == kBadRow) continue;
          stC((uint64_t)rowm[a] + co, cval(a, b, r));
        }
      }
  } else {
    uint32_t coln[FN];
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      const int n = n_of(b, 0);
      coln[b] = (n < N) ? decode_offset((uint32_t)n, nrank, it->ndims, it->c_nstr) : kBadRow;
    }
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m_of(a, r);
        if (m >= M) continue;
        const uint32_t ro = decode_offset((uint32_t)m, mrank, it->mdims, it->c_mstr);
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          if (coln[b] == kBadRow) continue;
          stC((uint64_t)ro + coln[b], cval(a, b, r));
        }
      }
  }
}

XcdSplit uniform_xcd_split(int total_tiles) {
  XcdSplit s;
  const int q = total_tiles >> 3, r = total_tiles & 7;
  for (int x = 0; x <= 8; ++x) s.start[x] = x * q + (x < r ? x : r);
  return s;
}

int tiles_for(int config, int M, int N) {
  const TileConfig& c = kTileConfigs[config];
  return ((M + c.bm - 1) / c.bm) * ((N + c.bn - 1) / c.bn);
}

template <int BM, int BN, int BK, int WM, int WN, bool M3, bool MAP>
static hipError_t launch_f64_map(const void* A, const void* B, void* C, const GemmItem* items,
                                 int nitems, const GemmSeg* segs, const uint32_t* koff,
                                 int total_tiles, int conjA, int conjB, int ksplit, void* partial,
                                 bool tr, FusedMaps maps, hipStream_t stream) {
  constexpr int NT = 64 * (BM / WM) * (BN / WN);
  if (tr)
    hipLaunchKernelGGL((zgemm_f64_kernel<BM, BN, BK, WM, WN, true, M3, MAP>),
                       dim3(total_tiles * ksplit), dim3(NT), 0, stream, (const double2*)A,
                       (const double2*)B, (double2*)C, items, nitems, segs, koff, total_tiles,
                       conjA, conjB, ksplit, (double2*)partial, maps);
  else
    hipLaunchKernelGGL((zgemm_f64_kernel<BM, BN, BK, WM, WN, false, M3, MAP>),
                       dim3(total_tiles * ksplit), dim3(NT), 0, stream, (const double2*)A,
                       (const double2*)B, (double2*)C, items, nitems, segs, koff, total_tiles,
                       conjA, conjB, ksplit, (double2*)partial, maps);
  return hipGetLastError();
}

template <int BM, int BN, int BK, int WM, int WN, bool M3 = false>
static hipError_t launch_f64(const void* A, const void* B, void* C, const GemmItem* items,
                             int nitems, const GemmSeg* segs, const uint32_t* koff,
                             int total_tiles, int conjA, int conjB, int ksplit, void* partial,
                             bool tr, const FusedMaps* maps, hipStream_t stream) {
  if (maps) {
    if (ksplit != 1) return hipErrorInvalidValue;   // mapped stores need whole tiles
    return launch_f64_map<BM, BN, BK, WM, WN, M3, true>(A, B, C, items, nitems, segs, koff,
                                                        total_tiles, conjA, conjB, 1, nullptr, tr,
                                                        *maps, stream);
  }
  return launch_f64_map<BM, BN, BK, WM, WN, M3, false>(A, B, C, items, nitems, segs, koff,
                                                       total_tiles, conjA, conjB, ksplit, partial,
                                                       tr, FusedMaps{}, stream);
}

hipError_t launch_zgemm_tiles(int config, const void* A, const void* B, void* C,
                              const GemmItem* items, int nitems, const GemmSeg* segs,
                              const uint32_t* koff, int total_tiles, int conjA, int conjB,
                              int ksplit, void* partial, bool tr, const FusedMaps* maps,
                              int kfast, const XcdSplit* xs, hipStream_t stream);

// Sum the k-chunk partials of a single-block item into its output layout.
__global__ void __launch_bounds__(256)
    splitk_reduce_kernel(const double2* __restrict__ partial, int ksplit,
                         const GemmItem* __restrict__ item, double2* __restrict__ Cbase) {
  const int M = item->M, N = item->N;
  const int64_t total = (int64_t)M * N;
  double2* C = Cbase + item->c_off;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int s = 0; s < ksplit; ++s) {
      const double2 v = partial[(size_t)s * total + e];
      re += v.x;
      im += v.y;
    }
    const int m = (int)(e % M), n = (int)(e / M);
    const uint32_t off = decode_offset((uint32_t)m, item->mrank, item->mdims, item->c_mstr) +
                         decode_offset((uint32_t)n, item->nrank, item->ndims, item->c_nstr);
    C[off] = double2{re, im};
  }
}

// C[e] = sum over s of slice s (total elements each): the chunk partials of
// a coupled step whose output blocks' pair lists were cut.
__global__ void __launch_bounds__(256)
    sum_slices_kernel(const double2* __restrict__ partial, int nslices, int64_t total,
                      double2* __restrict__ C) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int s = 0; s < nslices; ++s) {
      const double2 v = partial[(size_t)s * total + e];
      re += v.x;
      im += v.y;
    }
    C[e] = double2{re, im};
  }
}

hipError_t launch_sum_slices(const void* partial, int nslices, int64_t total, void* C,
                             hipStream_t stream) {
  if (total <= 0) return hipSuccess;
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(sum_slices_kernel, dim3(grid), dim3(256), 0, stream,
                     (const double2*)partial, nslices, total, (double2*)C);
  return hipGetLastError();
}

hipError_t launch_zgemm(int config, int precision, const void* A, const void* B, void* C,
                        const GemmItem* items, int nitems, const GemmSeg* segs,
                        const uint32_t* koff, int total_tiles, int conjA, int conjB,
                        int ksplit, void* partial, bool tr, hipStream_t stream,
                        const FusedMaps* maps, int kfast, const XcdSplit* xs) {
  if (total_tiles <= 0) return hipSuccess;
  if (precision != GRAFT_PREC_FP64 && precision != GRAFT_PREC_FP32 &&
      precision != GRAFT_PREC_BF16X3)
    return hipErrorNotSupported;
  if (maps && precision != GRAFT_PREC_FP64) return hipErrorInvalidValue;
  auto tiles = [&](int ks, void* part) {
    if (precision == GRAFT_PREC_FP64)
      return launch_zgemm_tiles(config, A, B, C, items, nitems, segs, koff, total_tiles, conjA,
                                conjB, ks, part, tr, maps, ks == 1 ? kfast : -1, ks == 1 ? xs : nullptr,
                                stream);
    return launch_zgemm_lowp(config, precision, A, B, C, items, nitems, segs, koff, total_tiles,
                             conjA, conjB, ks, part, tr, stream);
  };
  if (ksplit > 1) {
    if (nitems != 1 || partial == nullptr) return hipErrorInvalidValue;
    hipError_t e = tiles(ksplit, partial);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(1024), dim3(256), 0, stream,
                       (const double2*)partial, ksplit, items, (double2*)C);
    return hipGetLastError();
  }
  return tiles(1, nullptr);
}

hipError_t launch_zgemm_tiles(int config, const void* A, const void* B, void* C,
                              const GemmItem* items, int nitems, const GemmSeg* segs,
                              const uint32_t* koff, int total_tiles, int conjA, int conjB,
                              int ksplit, void* partial, bool tr, const FusedMaps* maps,
                              int kfast, const XcdSplit* xs, hipStream_t stream) {
  switch (config) {
    case 0: return launch_f64<128, 128, 16, 64, 64>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 1: return launch_f64<128, 64, 16, 64, 32>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 2: return launch_f64<64, 64, 16, 32, 32>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 3: return launch_f64<256, 16, 16, 64, 16>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 4: return launch_f64<64, 128, 16, 32, 64>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 5: return launch_f64<32, 32, 16, 16, 16>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 6: return launch_f64<16, 256, 16, 16, 64>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 7: return launch_f64<64, 64, 16, 32, 32, true>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 8: return launch_f64<128, 64, 16, 64, 32, true>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 9: return launch_f64<64, 128, 16, 32, 64, true>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 10: return launch_f64<64, 256, 8, 64, 64>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, stream);
    case 11: case 12: case 13: case 14: case 15: case 16: case 17: case 18: {
      // the double-buffered tiles address operands directly; mapped or
      // split-K steps run the single-buffered kernel of the same geometry
      if (maps == nullptr && ksplit == 1)
        return launch_zgemm_db(config, A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, tr, kfast, xs, stream);
      const int g = config >= 15 ? config - 4 : config;
      const int same = g == 11 ? 7 : g == 13 ? 9 : 8;
      return launch_zgemm_tiles(same, A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, ksplit, partial, tr, maps, -1, nullptr, stream);
    }
    default: return hipErrorInvalidValue;
  }
}

}  // namespace graft
