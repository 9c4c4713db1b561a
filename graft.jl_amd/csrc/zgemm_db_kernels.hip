// Double-buffered three-product complex FP64 GEMM for dense and grouped
// steps (tile configurations 11-14).
//
// Same items / segments / permute-fused addressing as zgemm_f64_kernel, with
// a different pipeline:
//  - operands sit in LDS as complex values (one 16-byte entry per element,
//    conjugation applied when stored), so a tile stage is one ds_write_b128
//    per element and an MFMA fragment one ds_read_b128 per lane; the sums
//    Ar + Ai and Br + Bi of the three-product form are formed in registers;
//  - two LDS buffers: the global loads of k-tile t + 1 are in flight while
//    tile t is multiplied out of one buffer, then stored into the other, and
//    a single barrier per k-tile separates the two.
// The three real products per 16x16 block and 4 complex k values are
//     T1 = Ar Br,  T2 = Ai Bi,  T3 = (Ar + Ai)(Br + Bi),
//     Cr = T1 - T2,  Ci = T3 - T1 - T2.
#include <type_traits>

#include "heff_internal.h"

namespace graft {

typedef double dbl4 __attribute__((ext_vector_type(4)));
// Staging registers hold native two-lane vectors: with HIP's double2 struct,
// a load into a register array followed by a store of it is folded into one
// global-to-LDS copy, which moves the load down to the store.
typedef double dbl2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ dbl2 ld2(const double2* p) { return *(const dbl2*)p; }
__device__ __forceinline__ void st2(double2* p, dbl2 v) { *(dbl2*)p = v; }

static constexpr uint32_t kBadRowDb = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t decode_offset_db(uint32_t idx, int rank, const int32_t* dims,
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

// Item of a tile: the last item whose first tile is <= tile.  Up to 256
// items are searched in LDS after one cooperative load of their first
// tiles (one memory round trip instead of one per bisection step).
template <int NT>
__device__ __forceinline__ int find_item(const GemmItem* __restrict__ items, int nitems, int tile) {
  if (nitems == 1) return 0;
  int lo = 0, hi = nitems - 1;
  if (nitems <= 256) {
    __shared__ int tb[256];
    for (int i = threadIdx.x; i < nitems; i += NT) tb[i] = items[i].tile_begin;
    __syncthreads();
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tb[mid] <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
  }
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <int BM, int BN, int BK, int WM, int WN, bool TR>
__global__ void __launch_bounds__(64 * (BM / WM) * (BN / WN))
    zgemm3m_db_kernel(const double2* __restrict__ Abase, const double2* __restrict__ Bbase,
                      double2* __restrict__ Cbase, const GemmItem* __restrict__ items, int nitems,
                      const GemmSeg* __restrict__ segs, const uint32_t* __restrict__ koff,
                      int total_tiles, int conjA, int conjB, XcdSplit xs) {
  constexpr int NWN = BN / WN;
  constexpr int NT = 64 * (BM / WM) * NWN;
  constexpr int EA = BM * BK / NT;
  constexpr int EB = BN * BK / NT;
  constexpr int FM = WM / 16;
  constexpr int FN = WN / 16;
  constexpr int LA = BM + 1;   // complex entries per k row of a stage (+1 staggers the banks)
  constexpr int LB = BN + 1;
  static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0, "tile/thread mismatch");
  static_assert(EA >= 1 && EB >= 1 && BK % 4 == 0, "tile too small for the thread count");

  __shared__ double2 As[2][BK * LA];
  __shared__ double2 Bs[2][BK * LB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / NWN) * WM;
  const int wn0 = (wave % NWN) * WN;
  const int li = lane & 15;
  const int kr = lane >> 4;

  // XCD-aware tile order (block b runs on XCD b % 8): each XCD walks a
  // contiguous range of logical tiles, so tiles sharing an A panel share an L2.
  // work-balanced XCD ranges (blocks past their XCD's range exit at once)
  const int tile = xs.start[blockIdx.x & 7] + (int)(blockIdx.x >> 3);
  if (tile >= xs.start[(blockIdx.x & 7) + 1]) return;
  const GemmItem* it = items + find_item<64 * (BM / WM) * (BN / WN)>(items, nitems, tile);
  const int M = it->M, N = it->N;
  const int mrank = it->mrank, nrank = it->nrank;
  const int local = tile - it->tile_begin;
  const int m0 = (local / it->tiles_n) * BM;
  const int n0 = (local % it->tiles_n) * BN;

  dbl4 acc[3][FM][FN];
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) acc[t][a][b] = dbl4{0.0, 0.0, 0.0, 0.0};

  const double sA = conjA ? -1.0 : 1.0;
  const double sB = conjB ? -1.0 : 1.0;

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

    // Row / column offsets of this thread's elements.  Rows past M and k
    // past K read a valid element (row 0, k clamped to K - 1) that is then
    // replaced by zero, so every load is unconditional.
    uint32_t rowA[EA], colB[EB];
    uint32_t okA = 0, okB = 0;   // bit e: element e's row / column exists
    if (akf) {
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int m = m0 + tid / BK + e * (NT / BK);
        const bool ok = m < M;
        rowA[e] = ok ? decode_offset_db((uint32_t)m, mrank, it->mdims, sg->a_mstr) : 0u;
        okA |= (uint32_t)ok << e;
      }
    } else {
      const int m = m0 + tid % BM;
      const bool ok = m < M;
      rowA[0] = ok ? decode_offset_db((uint32_t)m, mrank, it->mdims, sg->a_mstr) : 0u;
      okA = ok ? ~0u : 0u;
    }
    if (bkf) {
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int n = n0 + tid / BK + e * (NT / BK);
        const bool ok = n < N;
        colB[e] = ok ? decode_offset_db((uint32_t)n, nrank, it->ndims, sg->b_nstr) : 0u;
        okB |= (uint32_t)ok << e;
      }
    } else {
      const int n = n0 + tid % BN;
      const bool ok = n < N;
      colB[0] = ok ? decode_offset_db((uint32_t)n, nrank, it->ndims, sg->b_nstr) : 0u;
      okB = ok ? ~0u : 0u;
    }

    // Contracted offsets are read from the step's table one k-tile ahead of
    // the data loads that use them, so no data load waits on a table load.
    uint32_t ka[EA], kb[EB];
    auto loadk = [&](int k0) {
      if (akf) {
        ka[0] = kA[min(k0 + tid % BK, K - 1)];
      } else {
#pragma unroll
        for (int e = 0; e < EA; ++e) ka[e] = kA[min(k0 + tid / BM + e * (NT / BM), K - 1)];
      }
      if (bkf) {
        kb[0] = kB[min(k0 + tid % BK, K - 1)];
      } else {
#pragma unroll
        for (int e = 0; e < EB; ++e) kb[e] = kB[min(k0 + tid / BN + e * (NT / BN), K - 1)];
      }
    };
    // fetch issues the loads of one k-tile; store writes them into a stage.
    // Interior k-tiles of interior tiles without conjugation (a uniform
    // test) store the loaded values as they are; the others record which
    // elements exist at fetch time and zero the rest while storing, so no
    // instruction between fetch and store waits on the loads.
    const bool plain = !conjA && !conjB && m0 + BM <= M && n0 + BN <= N;
    dbl2 ra[EA], rb[EB];
    uint32_t vA = 0, vB = 0;   // bit e: fetched element e exists
    auto fetch = [&](int k0) {
      if (akf) {
#pragma unroll
        for (int e = 0; e < EA; ++e) ra[e] = ld2(A + ((uint64_t)rowA[e] + ka[0]));
      } else {
#pragma unroll
        for (int e = 0; e < EA; ++e) ra[e] = ld2(A + ((uint64_t)rowA[0] + ka[e]));
      }
      if (bkf) {
#pragma unroll
        for (int e = 0; e < EB; ++e) rb[e] = ld2(B + ((uint64_t)colB[e] + kb[0]));
      } else {
#pragma unroll
        for (int e = 0; e < EB; ++e) rb[e] = ld2(B + ((uint64_t)colB[0] + kb[e]));
      }
      if (!(plain && k0 + BK <= K)) {
        if (akf) {
          vA = k0 + tid % BK < K ? okA : 0u;
        } else {
          uint32_t v = 0;
#pragma unroll
          for (int e = 0; e < EA; ++e) v |= (uint32_t)(k0 + tid / BM + e * (NT / BM) < K) << e;
          vA = okA ? v : 0u;
        }
        if (bkf) {
          vB = k0 + tid % BK < K ? okB : 0u;
        } else {
          uint32_t v = 0;
#pragma unroll
          for (int e = 0; e < EB; ++e) v |= (uint32_t)(k0 + tid / BN + e * (NT / BN) < K) << e;
          vB = okB ? v : 0u;
        }
      }
    };
    auto store = [&](int buf, int k0) {
      double2* as = As[buf];
      double2* bs = Bs[buf];
      if (plain && k0 + BK <= K) {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
          const int q = akf ? tid % BK : tid / BM + e * (NT / BM);
          const int p = akf ? tid / BK + e * (NT / BK) : tid % BM;
          st2(&as[q * LA + p], ra[e]);
        }
#pragma unroll
        for (int e = 0; e < EB; ++e) {
          const int q = bkf ? tid % BK : tid / BN + e * (NT / BN);
          const int p = bkf ? tid / BK + e * (NT / BK) : tid % BN;
          st2(&bs[q * LB + p], rb[e]);
        }
        return;
      }
#pragma unroll
      for (int e = 0; e < EA; ++e) {
        const int q = akf ? tid % BK : tid / BM + e * (NT / BM);
        const int p = akf ? tid / BK + e * (NT / BK) : tid % BM;
        const bool ok = (vA >> e) & 1u;
        st2(&as[q * LA + p], dbl2{ok ? ra[e].x : 0.0, ok ? sA * ra[e].y : 0.0});
      }
#pragma unroll
      for (int e = 0; e < EB; ++e) {
        const int q = bkf ? tid % BK : tid / BN + e * (NT / BN);
        const int p = bkf ? tid / BK + e * (NT / BK) : tid % BN;
        const bool ok = (vB >> e) & 1u;
        st2(&bs[q * LB + p], dbl2{ok ? rb[e].x : 0.0, ok ? sB * rb[e].y : 0.0});
      }
    };
    auto compute = [&](int buf) {
      const double2* as = As[buf] + wm0 + li;
      const double2* bs = Bs[buf] + wn0 + li;
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kq = kk + kr;   // lane group kr supplies complex k = kk + kr
        dbl2 va[FM], vb[FN];
        double sa[FM], sb[FN];
#pragma unroll
        for (int a = 0; a < FM; ++a) va[a] = ld2(as + kq * LA + a * 16);
#pragma unroll
        for (int b = 0; b < FN; ++b) vb[b] = ld2(bs + kq * LB + b * 16);
#pragma unroll
        for (int a = 0; a < FM; ++a) sa[a] = va[a].x + va[a].y;
#pragma unroll
        for (int b = 0; b < FN; ++b) sb[b] = vb[b].x + vb[b].y;
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            if constexpr (TR) {
              acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(vb[b].x, va[a].x, acc[0][a][b], 0, 0, 0);
              acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(vb[b].y, va[a].y, acc[1][a][b], 0, 0, 0);
              acc[2][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(sb[b], sa[a], acc[2][a][b], 0, 0, 0);
            } else {
              acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[a].x, vb[b].x, acc[0][a][b], 0, 0, 0);
              acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[a].y, vb[b].y, acc[1][a][b], 0, 0, 0);
              acc[2][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(sa[a], sb[b], acc[2][a][b], 0, 0, 0);
            }
          }
      }
    };

    // Pipeline: stage tile 0; then per k-tile t: loads of t + 1 in flight,
    // multiply t, store t + 1 into the other buffer, one barrier.  The
    // previous segment ended on a barrier after its last multiply, so
    // buffer 0 is free here.
    const int nkt = (K + BK - 1) / BK;
    if (nkt == 0) continue;
    loadk(0);
    fetch(0);
    if (nkt > 1) loadk(BK);
    store(0, 0);
    __syncthreads();
    for (int kt = 0; kt < nkt; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nkt) fetch((kt + 1) * BK);
      if (kt + 2 < nkt) loadk((kt + 2) * BK);
      compute(buf);
      if (kt + 1 < nkt) store(buf ^ 1, (kt + 1) * BK);
      __syncthreads();
    }
  }

  // Epilogue: v_mfma_f64_16x16x4f64 leaves D[kr + 4*r][li] in accumulator
  // element r of lane (16*kr + li); D is the (m, n) block, or (n, m) under TR.
  auto m_of = [&](int a, int r) { return m0 + wm0 + a * 16 + (TR ? li : kr + 4 * r); };
  auto n_of = [&](int b, int r) { return n0 + wn0 + b * 16 + (TR ? kr + 4 * r : li); };
  auto cval = [&](int a, int b, int r) -> double2 {
    const double t1 = acc[0][a][b][r], t2 = acc[1][a][b][r], t3 = acc[2][a][b][r];
    return double2{t1 - t2, t3 - t1 - t2};
  };
  double2* C = Cbase + it->c_off;
  if constexpr (TR) {
    uint32_t rowm[FM];
#pragma unroll
    for (int a = 0; a < FM; ++a) {
      const int m = m_of(a, 0);
      rowm[a] = (m < M) ? decode_offset_db((uint32_t)m, mrank, it->mdims, it->c_mstr) : kBadRowDb;
    }
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n_of(b, r);
        if (n >= N) continue;
        const uint32_t co = decode_offset_db((uint32_t)n, nrank, it->ndims, it->c_nstr);
#pragma unroll
        for (int a = 0; a < FM; ++a) {
          if (rowm[a] == kBadRowDb) continue;
          C[(uint64_t)rowm[a] + co] = cval(a, b, r);
        }
      }
  } else {
    uint32_t coln[FN];
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      const int n = n_of(b, 0);
      coln[b] = (n < N) ? decode_offset_db((uint32_t)n, nrank, it->ndims, it->c_nstr) : kBadRowDb;
    }
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m_of(a, r);
        if (m >= M) continue;
        const uint32_t ro = decode_offset_db((uint32_t)m, mrank, it->mdims, it->c_mstr);
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          if (coln[b] == kBadRowDb) continue;
          C[(uint64_t)ro + coln[b]] = cval(a, b, r);
        }
      }
  }
}

// Pipelined variant for steps whose every tile is whole (M % BM, N % BN and
// K % BK all zero), without conjugation, and whose segments all load A and B
// in one direction (AKF / BKF: along k).  No element needs masking, so a
// k-tile is straight-line code: the loads of k-tile t + 1 and the table
// reads of t + 2 are issued first, then the MFMAs of t with the fragments of
// step kk + 1 read from LDS into a second register set while those of kk
// issue, then the stage store and one barrier.
template <int BM, int BN, int BK, int WM, int WN, bool TR, bool AKF, bool BKF>
__global__ void __launch_bounds__(64 * (BM / WM) * (BN / WN))
    zgemm3m_piped_kernel(const double2* __restrict__ Abase, const double2* __restrict__ Bbase,
                         double2* __restrict__ Cbase, const GemmItem* __restrict__ items, int nitems,
                         const GemmSeg* __restrict__ segs, const uint32_t* __restrict__ koff,
                         int total_tiles, XcdSplit xs) {
  constexpr int NWN = BN / WN;
  constexpr int NT = 64 * (BM / WM) * NWN;
  constexpr int EA = BM * BK / NT;
  constexpr int EB = BN * BK / NT;
  constexpr int FM = WM / 16;
  constexpr int FN = WN / 16;
  constexpr int LA = BM + 1;
  constexpr int LB = BN + 1;
  constexpr int NKK = BK / 4;
  static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0, "tile/thread mismatch");
  static_assert(EA >= 1 && EB >= 1 && NKK % 2 == 0, "tile too small for the thread count");

  __shared__ double2 As[2][BK * LA];
  __shared__ double2 Bs[2][BK * LB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / NWN) * WM;
  const int wn0 = (wave % NWN) * WN;
  const int li = lane & 15;
  const int kr = lane >> 4;

  // work-balanced XCD ranges (blocks past their XCD's range exit at once)
  const int tile = xs.start[blockIdx.x & 7] + (int)(blockIdx.x >> 3);
  if (tile >= xs.start[(blockIdx.x & 7) + 1]) return;
  const GemmItem* it = items + find_item<64 * (BM / WM) * (BN / WN)>(items, nitems, tile);
  const int mrank = it->mrank, nrank = it->nrank;
  const int local = tile - it->tile_begin;
  const int m0 = (local / it->tiles_n) * BM;
  const int n0 = (local % it->tiles_n) * BN;

  // LDS positions of this thread's stage elements
  int sa_[EA], sb_[EB];
#pragma unroll
  for (int e = 0; e < EA; ++e)
    sa_[e] = AKF ? (tid % BK) * LA + tid / BK + e * (NT / BK) : (tid / BM + e * (NT / BM)) * LA + tid % BM;
#pragma unroll
  for (int e = 0; e < EB; ++e)
    sb_[e] = BKF ? (tid % BK) * LB + tid / BK + e * (NT / BK) : (tid / BN + e * (NT / BN)) * LB + tid % BN;

  dbl4 acc[3][FM][FN];
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) acc[t][a][b] = dbl4{0.0, 0.0, 0.0, 0.0};

  const int seg_end = it->seg_begin + it->seg_count;
  for (int s = it->seg_begin; s < seg_end; ++s) {
    const GemmSeg* sg = segs + s;
    const double2* A = Abase + sg->a_off;
    const double2* B = Bbase + sg->b_off;
    const int K = sg->K;
    const uint32_t* kA = koff + sg->koff_idx;
    const uint32_t* kB = kA + K;
    const int nkt = K / BK;

    uint32_t rowA[AKF ? EA : 1], colB[BKF ? EB : 1];
#pragma unroll
    for (int e = 0; e < (AKF ? EA : 1); ++e)
      rowA[e] = decode_offset_db((uint32_t)(m0 + (AKF ? tid / BK + e * (NT / BK) : tid % BM)), mrank,
                                 it->mdims, sg->a_mstr);
#pragma unroll
    for (int e = 0; e < (BKF ? EB : 1); ++e)
      colB[e] = decode_offset_db((uint32_t)(n0 + (BKF ? tid / BK + e * (NT / BK) : tid % BN)), nrank,
                                 it->ndims, sg->b_nstr);

    uint32_t pa[AKF ? 1 : EA], pb[BKF ? 1 : EB];
    auto koffs = [&](int k0) {
#pragma unroll
      for (int e = 0; e < (AKF ? 1 : EA); ++e)
        pa[e] = kA[min(k0 + (AKF ? tid % BK : tid / BM + e * (NT / BM)), K - 1)];
#pragma unroll
      for (int e = 0; e < (BKF ? 1 : EB); ++e)
        pb[e] = kB[min(k0 + (BKF ? tid % BK : tid / BN + e * (NT / BN)), K - 1)];
    };
    dbl2 ra[EA], rb[EB];
    auto loads = [&]() {
#pragma unroll
      for (int e = 0; e < EA; ++e)
        ra[e] = ld2(A + ((uint64_t)rowA[AKF ? e : 0] + pa[AKF ? 0 : e]));
#pragma unroll
      for (int e = 0; e < EB; ++e)
        rb[e] = ld2(B + ((uint64_t)colB[BKF ? e : 0] + pb[BKF ? 0 : e]));
    };
    auto stage = [&](int buf) {
#pragma unroll
      for (int e = 0; e < EA; ++e) st2(&As[buf][sa_[e]], ra[e]);
#pragma unroll
      for (int e = 0; e < EB; ++e) st2(&Bs[buf][sb_[e]], rb[e]);
    };
    dbl2 fa[2][FM], fb[2][FN];
    auto frag = [&](int set, int buf, int kk) {
      const double2* as = As[buf] + (kk * 4 + kr) * LA + wm0 + li;
      const double2* bs = Bs[buf] + (kk * 4 + kr) * LB + wn0 + li;
#pragma unroll
      for (int a = 0; a < FM; ++a) fa[set][a] = ld2(as + a * 16);
#pragma unroll
      for (int b = 0; b < FN; ++b) fb[set][b] = ld2(bs + b * 16);
    };
    auto mma = [&](int set) {
      double sa[FM], sb[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) sa[a] = fa[set][a].x + fa[set][a].y;
#pragma unroll
      for (int b = 0; b < FN; ++b) sb[b] = fb[set][b].x + fb[set][b].y;
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          if constexpr (TR) {
            acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[set][b].x, fa[set][a].x, acc[0][a][b], 0, 0, 0);
            acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[set][b].y, fa[set][a].y, acc[1][a][b], 0, 0, 0);
            acc[2][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(sb[b], sa[a], acc[2][a][b], 0, 0, 0);
          } else {
            acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[set][a].x, fb[set][b].x, acc[0][a][b], 0, 0, 0);
            acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[set][a].y, fb[set][b].y, acc[1][a][b], 0, 0, 0);
            acc[2][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(sa[a], sb[b], acc[2][a][b], 0, 0, 0);
          }
        }
    };
    auto tile_mma = [&](int buf) {
#pragma unroll
      for (int kk = 0; kk < NKK; ++kk) {
        if (kk + 1 < NKK) frag((kk + 1) & 1, buf, kk + 1);
        mma(kk & 1);
      }
    };

    // The previous segment ended on a barrier after its last multiply.  The
    // last k-tile also loads and stages a (clamped, unused) next tile, which
    // keeps the loop body free of branches.
    koffs(0);
    loads();
    koffs(BK);
    stage(0);
    __syncthreads();
    frag(0, 0, 0);
    // The compiler and scheduling barriers keep the loads ahead of the
    // MFMAs and the stage store behind them, where the loads have long
    // completed.
    for (int kt = 0; kt < nkt; ++kt) {
      const int buf = kt & 1;
      loads();                  // k-tile kt + 1
      koffs((kt + 2) * BK);     // k-tile kt + 2
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      tile_mma(buf);
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      stage(buf ^ 1);
      __syncthreads();
      frag(0, buf ^ 1, 0);
    }
  }

  auto m_of = [&](int a, int r) { return m0 + wm0 + a * 16 + (TR ? li : kr + 4 * r); };
  auto n_of = [&](int b, int r) { return n0 + wn0 + b * 16 + (TR ? kr + 4 * r : li); };
  auto cval = [&](int a, int b, int r) -> double2 {
    const double t1 = acc[0][a][b][r], t2 = acc[1][a][b][r], t3 = acc[2][a][b][r];
    return double2{t1 - t2, t3 - t1 - t2};
  };
  double2* C = Cbase + it->c_off;
  if constexpr (TR) {
    uint32_t rowm[FM];
#pragma unroll
    for (int a = 0; a < FM; ++a) rowm[a] = decode_offset_db((uint32_t)m_of(a, 0), mrank, it->mdims, it->c_mstr);
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t co = decode_offset_db((uint32_t)n_of(b, r), nrank, it->ndims, it->c_nstr);
#pragma unroll
        for (int a = 0; a < FM; ++a) C[(uint64_t)rowm[a] + co] = cval(a, b, r);
      }
  } else {
    uint32_t coln[FN];
#pragma unroll
    for (int b = 0; b < FN; ++b) coln[b] = decode_offset_db((uint32_t)n_of(b, 0), nrank, it->ndims, it->c_nstr);
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t ro = decode_offset_db((uint32_t)m_of(a, r), mrank, it->mdims, it->c_mstr);
#pragma unroll
        for (int b = 0; b < FN; ++b) C[(uint64_t)ro + coln[b]] = cval(a, b, r);
      }
  }
}

// Fused sector steps (per-charge matrices, one segment per item, rank-1
// row and column legs, contracted index linear in the fused layout) with
// element maps: an operand with a map is read through it from its blocked
// tensor, the others from their gathered fused matrices, and the result is
// stored through the C map.  Same double-buffered pipeline; every element's
// source index (map value, or fused index) is read one k-tile ahead of its
// data load, and the C map values of the tile are read before the k loop.
template <int BM, int BN, int BK, int WM, int WN, bool TR, bool AKF, bool BKF, bool PIPE>
__global__ void __launch_bounds__(64 * (BM / WM) * (BN / WN), 2)
    zgemm3m_map_kernel(const double2* __restrict__ Abase, const double2* __restrict__ Bbase,
                       double2* __restrict__ Cbase, const GemmItem* __restrict__ items, int nitems,
                       const GemmSeg* __restrict__ segs, int total_tiles, int conjA, int conjB,
                       FusedMaps maps, XcdSplit xs) {
  constexpr int NWN = BN / WN;
  constexpr int NT = 64 * (BM / WM) * NWN;
  constexpr int EA = BM * BK / NT;
  constexpr int EB = BN * BK / NT;
  constexpr int FM = WM / 16;
  constexpr int FN = WN / 16;
  constexpr int LA = BM + 1;
  constexpr int LB = BN + 1;
  static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0, "tile/thread mismatch");
  static_assert(EA >= 1 && EB >= 1 && BK % 4 == 0, "tile too small for the thread count");

  __shared__ double2 As[2][BK * LA];
  __shared__ double2 Bs[2][BK * LB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / NWN) * WM;
  const int wn0 = (wave % NWN) * WN;
  const int li = lane & 15;
  const int kr = lane >> 4;

  // work-balanced XCD ranges (blocks past their XCD's range exit at once)
  const int tile = xs.start[blockIdx.x & 7] + (int)(blockIdx.x >> 3);
  if (tile >= xs.start[(blockIdx.x & 7) + 1]) return;
  const GemmItem* it = items + find_item<64 * (BM / WM) * (BN / WN)>(items, nitems, tile);
  const int M = it->M, N = it->N;
  const int local = tile - it->tile_begin;
  const int m0 = (local / it->tiles_n) * BM;
  const int n0 = (local % it->tiles_n) * BN;
  const bool mapa = maps.a != nullptr, mapb = maps.b != nullptr, mapc = maps.c != nullptr;

  auto m_of = [&](int a, int r) { return m0 + wm0 + a * 16 + (TR ? li : kr + 4 * r); };
  auto n_of = [&](int b, int r) { return n0 + wn0 + b * 16 + (TR ? kr + 4 * r : li); };
  // destination of every output element of this thread (kMapZero: not
  // stored), read before the k loop so the epilogue does not wait on it
  const uint32_t* mC = mapc ? maps.c + it->c_off : nullptr;
  uint32_t cdst[FM][FN][4];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m_of(a, r), n = n_of(b, r);
        const bool ok = m < M && n < N;
        const uint32_t ci = ok ? (uint32_t)m * it->c_mstr[0] + (uint32_t)n * it->c_nstr[0] : 0u;
        cdst[a][b][r] = !ok ? kMapZero : mapc ? mC[ci] : (uint32_t)it->c_off + ci;
      }
  dbl4 acc[3][FM][FN];
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) acc[t][a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
  const double sA = conjA ? -1.0 : 1.0;
  const double sB = conjB ? -1.0 : 1.0;

  const GemmSeg* sg = segs + it->seg_begin;
  const int K = sg->K;
  const double2* A = mapa ? Abase : Abase + sg->a_off;
  const double2* B = mapb ? Bbase : Bbase + sg->b_off;
  const uint32_t* mA = mapa ? maps.a + sg->a_off : nullptr;
  const uint32_t* mB = mapb ? maps.b + sg->b_off : nullptr;
  const uint32_t akstr = AKF ? 1u : (uint32_t)M;    // fused A: M x K, k stride
  const uint32_t bkstr = BKF ? 1u : (uint32_t)N;    // fused B: K x N
  // row / column and k offset (within a tile) of this thread's stage element e
  const int pa0 = AKF ? tid / BK : tid % BM, qa0 = AKF ? tid % BK : tid / BM;
  const int pb0 = BKF ? tid / BK : tid % BN, qb0 = BKF ? tid % BK : tid / BN;
  auto pA = [&](int e) { return AKF ? pa0 + e * (NT / BK) : pa0; };
  auto qA = [&](int e) { return AKF ? qa0 : qa0 + e * (NT / BM); };
  auto pB = [&](int e) { return BKF ? pb0 + e * (NT / BK) : pb0; };
  auto qB = [&](int e) { return BKF ? qb0 : qb0 + e * (NT / BN); };
  uint32_t sa[EA], sb[EB];    // source indices of the next fetch (kMapZero: zero)
  auto sources = [&](int k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int m = m0 + pA(e), k = k0 + qA(e);
      const bool ok = m < M && k < K;
      const uint32_t fi = ok ? (uint32_t)m * sg->a_mstr[0] + (uint32_t)k * akstr : 0u;
      sa[e] = !ok ? kMapZero : mapa ? mA[fi] : fi;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int n = n0 + pB(e), k = k0 + qB(e);
      const bool ok = n < N && k < K;
      const uint32_t fi = ok ? (uint32_t)n * sg->b_nstr[0] + (uint32_t)k * bkstr : 0u;
      sb[e] = !ok ? kMapZero : mapb ? mB[fi] : fi;
    }
  };
  dbl2 ra[EA], rb[EB];
  uint32_t vA = 0, vB = 0;
  auto fetch = [&]() {
    vA = vB = 0;
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const bool ok = sa[e] < kMapZero;
      ra[e] = ld2(A + (ok ? sa[e] : 0u));
      vA |= (uint32_t)ok << e;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const bool ok = sb[e] < kMapZero;
      rb[e] = ld2(B + (ok ? sb[e] : 0u));
      vB |= (uint32_t)ok << e;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const bool ok = (vA >> e) & 1u;
      st2(&As[buf][qA(e) * LA + pA(e)], dbl2{ok ? ra[e].x : 0.0, ok ? sA * ra[e].y : 0.0});
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const bool ok = (vB >> e) & 1u;
      st2(&Bs[buf][qB(e) * LB + pB(e)], dbl2{ok ? rb[e].x : 0.0, ok ? sB * rb[e].y : 0.0});
    }
  };
  auto compute = [&](int buf) {
    const double2* as = As[buf] + wm0 + li;
    const double2* bs = Bs[buf] + wn0 + li;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kq = kk + kr;
      dbl2 va[FM], vb[FN];
      double su[FM], sv[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) va[a] = ld2(as + kq * LA + a * 16);
#pragma unroll
      for (int b = 0; b < FN; ++b) vb[b] = ld2(bs + kq * LB + b * 16);
#pragma unroll
      for (int a = 0; a < FM; ++a) su[a] = va[a].x + va[a].y;
#pragma unroll
      for (int b = 0; b < FN; ++b) sv[b] = vb[b].x + vb[b].y;
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          if constexpr (TR) {
            acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(vb[b].x, va[a].x, acc[0][a][b], 0, 0, 0);
            acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(vb[b].y, va[a].y, acc[1][a][b], 0, 0, 0);
            acc[2][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(sv[b], su[a], acc[2][a][b], 0, 0, 0);
          } else {
            acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[a].x, vb[b].x, acc[0][a][b], 0, 0, 0);
            acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[a].y, vb[b].y, acc[1][a][b], 0, 0, 0);
            acc[2][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(su[a], sv[b], acc[2][a][b], 0, 0, 0);
          }
        }
    }
  };

  const int nkt = (K + BK - 1) / BK;
  sources(0);
  fetch();
  sources(BK);
  store(0);
  __syncthreads();
  if constexpr (!PIPE) {
    for (int kt = 0; kt < nkt; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nkt) {
        fetch();                    // k-tile kt + 1
        sources((kt + 2) * BK);     // k-tile kt + 2 (out of range: zero)
      }
      compute(buf);
      if (kt + 1 < nkt) store(buf ^ 1);
      __syncthreads();
    }
  } else {
    // Pipelined form (the dense pipelined kernel's schedule): the MFMA
    // fragments are double-buffered in registers, so the LDS reads of step
    // kk + 1 are in flight while the MFMAs of step kk issue; the next
    // k-tile's global loads go out before the MFMA block and are staged
    // after it; the loop body has no branches.  The last k-tile runs outside
    // the loop (nothing left to load) and only up to K.
    constexpr int NKK = BK / 4;
    static_assert(NKK % 2 == 0, "fragment double-buffering needs an even k-step count");
    dbl2 fa[2][FM], fb[2][FN];
    auto frag = [&](int set, int buf, int kk) {
      const double2* as = As[buf] + (kk * 4 + kr) * LA + wm0 + li;
      const double2* bs = Bs[buf] + (kk * 4 + kr) * LB + wn0 + li;
#pragma unroll
      for (int a = 0; a < FM; ++a) fa[set][a] = ld2(as + a * 16);
#pragma unroll
      for (int b = 0; b < FN; ++b) fb[set][b] = ld2(bs + b * 16);
    };
    auto mma = [&](int set) {
      double su[FM], sv[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) su[a] = fa[set][a].x + fa[set][a].y;
#pragma unroll
      for (int b = 0; b < FN; ++b) sv[b] = fb[set][b].x + fb[set][b].y;
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          if constexpr (TR) {
            acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[set][b].x, fa[set][a].x, acc[0][a][b], 0, 0, 0);
            acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[set][b].y, fa[set][a].y, acc[1][a][b], 0, 0, 0);
            acc[2][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(sv[b], su[a], acc[2][a][b], 0, 0, 0);
          } else {
            acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[set][a].x, fb[set][b].x, acc[0][a][b], 0, 0, 0);
            acc[1][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[set][a].y, fb[set][b].y, acc[1][a][b], 0, 0, 0);
            acc[2][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(su[a], sv[b], acc[2][a][b], 0, 0, 0);
          }
        }
    };
    frag(0, 0, 0);
    for (int kt = 0; kt + 1 < nkt; ++kt) {
      const int buf = kt & 1;
      fetch();                    // k-tile kt + 1
      sources((kt + 2) * BK);     // k-tile kt + 2 (zeros past K)
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int kk = 0; kk < NKK; ++kk) {
        if (kk + 1 < NKK) frag((kk + 1) & 1, buf, kk + 1);
        mma(kk & 1);
      }
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      store(buf ^ 1);
      __syncthreads();
      frag(0, buf ^ 1, 0);
    }
    // The last k-tile, peeled: its 4-deep k steps wholly past K hold zeros
    // and are skipped (components have ragged K).
    {
      const int nk = (K - (nkt - 1) * BK + 3) / 4;   // live k steps, 1 .. NKK
#pragma unroll
      for (int kk = 0; kk < NKK; ++kk) {
        if (kk >= nk) break;
        if (kk + 1 < NKK && kk + 1 < nk) frag((kk + 1) & 1, (nkt - 1) & 1, kk + 1);
        mma(kk & 1);
      }
    }
  }


#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t d = cdst[a][b][r];
        if (d >= kMapZero) continue;
        const double t1 = acc[0][a][b][r], t2 = acc[1][a][b][r], t3 = acc[2][a][b][r];
        st2(Cbase + d, dbl2{t1 - t2, t3 - t1 - t2});
      }
}

hipError_t launch_zgemm_dbmap(int config, const void* A, const void* B, void* C, const GemmItem* items,
                              int nitems, const GemmSeg* segs, int total_tiles, int conjA, int conjB,
                              bool tr, int kfast, const FusedMaps& maps, const XcdSplit* xsp,
                              hipStream_t stream) {
  if (config != 11 && config != 15 && config != 14 && config != 18) return hipErrorInvalidValue;
  const bool wide = config == 14 || config == 18;   // 128 x 64, 8 waves
  const bool piped = config == 15 || config == 18;
  const XcdSplit xs = xsp ? *xsp : uniform_xcd_split(total_tiles);
  const dim3 grid(xcd_grid(xs)), block(wide ? 512 : 256);
  const double2* a = (const double2*)A;
  const double2* b = (const double2*)B;
  double2* c = (double2*)C;
#define GRAFT_MAPK1(BMV, TRV, AKV, BKV, PV)                                                          \
  hipLaunchKernelGGL((zgemm3m_map_kernel<BMV, 64, 16, 32, 32, TRV, AKV, BKV, PV>), grid, block, 0,  \
                     stream, a, b, c, items, nitems, segs, total_tiles, conjA, conjB, maps, xs)
#define GRAFT_MAPK(TRV, AKV, BKV)                                                                   \
  if (wide) {                                                                                       \
    if (piped) GRAFT_MAPK1(128, TRV, AKV, BKV, true); else GRAFT_MAPK1(128, TRV, AKV, BKV, false);  \
  } else {                                                                                          \
    if (piped) GRAFT_MAPK1(64, TRV, AKV, BKV, true); else GRAFT_MAPK1(64, TRV, AKV, BKV, false);    \
  }
  switch ((kfast & 3) | (tr ? 4 : 0)) {
    case 0: GRAFT_MAPK(false, false, false); break;
    case 1: GRAFT_MAPK(false, true, false); break;
    case 2: GRAFT_MAPK(false, false, true); break;
    case 3: GRAFT_MAPK(false, true, true); break;
    case 4: GRAFT_MAPK(true, false, false); break;
    case 5: GRAFT_MAPK(true, true, false); break;
    case 6: GRAFT_MAPK(true, false, true); break;
    default: GRAFT_MAPK(true, true, true); break;
  }
#undef GRAFT_MAPK
#undef GRAFT_MAPK1
  return hipGetLastError();
}

template <int BM, int BN, int BK, int WM, int WN>
static hipError_t launch_db(const void* A, const void* B, void* C, const GemmItem* items,
                            int nitems, const GemmSeg* segs, const uint32_t* koff, int total_tiles,
                            int conjA, int conjB, bool tr, int kfast, const XcdSplit* xsp,
                            hipStream_t stream) {
  constexpr int NT = 64 * (BM / WM) * (BN / WN);
  const XcdSplit xs = xsp ? *xsp : uniform_xcd_split(total_tiles);
  const dim3 grid(xcd_grid(xs)), block(NT);
  const double2* a = (const double2*)A;
  const double2* b = (const double2*)B;
  double2* c = (double2*)C;
  if (kfast >= 0 && !conjA && !conjB) {
#define GRAFT_PIPED(TRV, AKV, BKV)                                                                  \
  hipLaunchKernelGGL((zgemm3m_piped_kernel<BM, BN, BK, WM, WN, TRV, AKV, BKV>), grid, block, 0,     \
                     stream, a, b, c, items, nitems, segs, koff, total_tiles, xs)
    switch (kfast | (tr ? 4 : 0)) {
      case 0: GRAFT_PIPED(false, false, false); break;
      case 1: GRAFT_PIPED(false, true, false); break;
      case 2: GRAFT_PIPED(false, false, true); break;
      case 3: GRAFT_PIPED(false, true, true); break;
      case 4: GRAFT_PIPED(true, false, false); break;
      case 5: GRAFT_PIPED(true, true, false); break;
      case 6: GRAFT_PIPED(true, false, true); break;
      default: GRAFT_PIPED(true, true, true); break;
    }
#undef GRAFT_PIPED
    return hipGetLastError();
  }
  if (tr)
    hipLaunchKernelGGL((zgemm3m_db_kernel<BM, BN, BK, WM, WN, true>), grid, block, 0, stream, a, b,
                       c, items, nitems, segs, koff, total_tiles, conjA, conjB, xs);
  else
    hipLaunchKernelGGL((zgemm3m_db_kernel<BM, BN, BK, WM, WN, false>), grid, block, 0, stream, a, b,
                       c, items, nitems, segs, koff, total_tiles, conjA, conjB, xs);
  return hipGetLastError();
}

hipError_t launch_zgemm_db(int config, const void* A, const void* B, void* C, const GemmItem* items,
                           int nitems, const GemmSeg* segs, const uint32_t* koff, int total_tiles,
                           int conjA, int conjB, bool tr, int kfast, const XcdSplit* xs,
                           hipStream_t stream) {
  const int kf = config >= 15 ? kfast : -1;
  switch (config) {
    case 11: case 15: return launch_db<64, 64, 16, 32, 32>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, tr, kf, xs, stream);
    case 12: case 16: return launch_db<128, 64, 16, 64, 32>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, tr, kf, xs, stream);
    case 13: case 17: return launch_db<64, 128, 16, 32, 64>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, tr, kf, xs, stream);
    case 14: case 18: return launch_db<128, 64, 16, 32, 32>(A, B, C, items, nitems, segs, koff, total_tiles, conjA, conjB, tr, kf, xs, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace graft
