// Internal structures shared by the runtime (heff_runtime.hip) and the
// kernels (zgemm_kernels.hip, vector_kernels.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/graft_heff.h"

namespace graft {

// Maximum number of merged legs in one GEMM index group (m, n or k).
constexpr int GL = 8;

// One contribution to an output block: C_blk += op(A_blk) * op(B_blk),
// contracted over K = prod(k dims).  Offsets are in complex elements.
struct GemmSeg {
  int64_t a_off;            // A block start inside the A buffer
  int64_t b_off;            // B block start inside the B buffer
  int32_t K;                // contracted extent
  int32_t koff_idx;         // koff[koff_idx + q] = A offset, koff[koff_idx + K + q] = B offset
  int32_t a_kfast;          // A's unit-stride leg is contracted (load along k)
  int32_t b_kfast;          // B's unit-stride leg is contracted (load along k)
  uint32_t a_mstr[GL];      // A element strides of the item's m legs
  uint32_t b_nstr[GL];      // B element strides of the item's n legs
};

// One output block and its tile grid.  m legs come from A, n legs from B.
struct GemmItem {
  int64_t c_off;
  int32_t M, N;
  int32_t mrank, nrank;
  int32_t mdims[GL];
  int32_t ndims[GL];
  uint32_t c_mstr[GL];
  uint32_t c_nstr[GL];
  int32_t seg_begin, seg_count;
  int32_t tile_begin, tiles_n;
};

// Tile configurations of the complex-FP64 MFMA GEMM.
struct TileConfig {
  int bm, bn, bk, wm, wn;
  int m3;  // three real products per complex product (Karatsuba form)
};
constexpr int kNumTileConfigs = 19;
extern const TileConfig kTileConfigs[kNumTileConfigs];

// Logical tiles of XCD x: [start[x], start[x + 1]).  Hardware dispatches
// block b to XCD b % 8, so the grid is 8 * max range and block b runs tile
// start[b % 8] + b / 8 when that is inside its XCD's range.  Ranges of equal
// work (k-tiles plus a fixed cost per tile) keep XCDs busy equally on steps
// whose items differ in K; contiguous ranges keep tiles sharing an operand
// panel on one L2.
struct XcdSplit {
  int32_t start[9];
};
XcdSplit uniform_xcd_split(int total_tiles);
inline int xcd_grid(const XcdSplit& s) {
  int g = 0;
  for (int x = 0; x < 8; ++x) g = g > s.start[x + 1] - s.start[x] ? g : s.start[x + 1] - s.start[x];
  return 8 * g;
}

// Element maps of a fused sector step read / written inside the GEMM: entry
// i of a fused matrix lives at blocked element a[i] / b[i] / c[i]
// (kMapZero: absent — zero when read, skipped when stored).
// coef (coupled steps, single-buffered tiles only): a real factor per
// segment, applied to A's values as they are staged — the recoupling
// coefficient of a non-abelian block pair.
struct FusedMaps {
  const uint32_t* a;
  const uint32_t* b;
  const uint32_t* c;
  const double* coef = nullptr;
};

// Launch one grouped strided complex GEMM (precision GRAFT_PREC_*).
// ksplit > 1 (single-block items only) splits K over ksplit blocks per
// tile; partial must hold ksplit * M * N complex values.  tr selects the
// transposed MFMA orientation (coalesced stores when C's unit-stride leg is
// an m leg).  maps (FP64, ksplit 1 only): the item/segment offsets address
// fused matrices and A, B, C are the blocked tensors behind the maps.
hipError_t launch_zgemm(int config, int precision, const void* A, const void* B, void* C,
                        const GemmItem* items, int nitems, const GemmSeg* segs,
                        const uint32_t* koff, int total_tiles, int conjA, int conjB,
                        int ksplit, void* partial, bool tr, hipStream_t stream,
                        const FusedMaps* maps = nullptr, int kfast = -1,
                        const XcdSplit* xs = nullptr);
// C[e] = sum of nslices consecutive slices of ``total`` complex elements.
hipError_t launch_sum_slices(const void* partial, int nslices, int64_t total, void* C,
                             hipStream_t stream);
int tiles_for(int config, int M, int N);
// FP32 / BF16x3 products with the same addressing (zgemm_lowp_kernels.hip).
// Double-buffered three-product FP64 tiles (configurations 11-18), operands
// addressed directly: no element maps, no split-K.  Configurations 15-18 run
// the pipelined kernel when kfast >= 0 (every tile whole, no conjugation,
// bit 0 / bit 1: every segment loads A / B along k) and the kernel of
// configuration c - 4 otherwise.
hipError_t launch_zgemm_db(int config, const void* A, const void* B, void* C, const GemmItem* items,
                           int nitems, const GemmSeg* segs, const uint32_t* koff, int total_tiles,
                           int conjA, int conjB, bool tr, int kfast, const XcdSplit* xs,
                           hipStream_t stream);
// Fused sector steps through element maps on the double-buffered tiles:
// 64x64 (configurations 11, 15) and 128x64 with 8 waves (14, 18); 15 and 18
// run the pipelined schedule.  kfast: bit 0 / bit 1 A / B fused matrices
// row-major (loaded along k).
hipError_t launch_zgemm_dbmap(int config, const void* A, const void* B, void* C, const GemmItem* items,
                              int nitems, const GemmSeg* segs, int total_tiles, int conjA, int conjB,
                              bool tr, int kfast, const FusedMaps& maps, const XcdSplit* xs,
                              hipStream_t stream);
// planes (BF16X3, ksplit 1): A holds split-bf16 word pairs (launch_zskinny
// planes) and B split-bf16 word quads (launch_split_bplanes) at the element
// indices of the FP64 layout; the tiles stage them without conversion.
hipError_t launch_zgemm_lowp(int config, int precision, const void* A, const void* B, void* C,
                             const GemmItem* items, int nitems, const GemmSeg* segs,
                             const uint32_t* koff, int total_tiles, int conjA, int conjB,
                             int ksplit, void* partial, bool tr, hipStream_t stream,
                             bool planes = false);

// Streaming kernel for skinny steps (skinny_kernels.hip).  It is selected as
// the virtual tile configuration kSkinnyConfig when one operand's open
// extent is <= kSkinnyMaxS and K <= kSkinnyMaxK.
constexpr int kSkinnyConfig = kNumTileConfigs;
constexpr int kSkinnyMaxK = 16;
constexpr int kSkinnyMaxS = 64;
constexpr int kSkinnyAccMaxK = 64;   // accumulator variant: longer K, smaller side
constexpr int kSkinnyAccMaxS = 16;
// acc selects the accumulator variant (K > kSkinnyMaxK or S > kSkinnyMaxS).
// planes: the result is written as split-bf16 word pairs (8 bytes per element,
// at the element's index) for a BF16x3 GEMM that reads it as its A operand.
hipError_t launch_zskinny(bool big_is_a, bool acc, const void* A, const void* B, void* C,
                          const GemmItem* item, const GemmSeg* seg, const uint32_t* koff, int grid,
                          int conjA, int conjB, hipStream_t stream, bool planes = false);
// Split-bf16 words of a complex FP64 B operand (16 bytes per element, four
// words: the real- and imaginary-part columns of the four-product form).
hipError_t launch_split_bplanes(const void* B, int64_t n, void* out, int conjB, hipStream_t stream);
int skinny_grid(int64_t J);

// Per-block streaming kernel for multi-block (sector) steps whose small
// operand has few open elements per output block: it reads the blocks in
// place (no gather into fused matrices) through the same item/segment/koff
// description.  Virtual tile configuration kBlockStreamConfig.
constexpr int kBlockStreamConfig = kNumTileConfigs + 1;
constexpr int kBlockStreamS = 16;       // small-side outputs per pass (registers per lane)
constexpr int kBlockStreamLaneRows = 1; // big-side elements per lane, one after another
constexpr int kBlockStreamRows = 64 * kBlockStreamLaneRows;   // big-side elements per work entry
constexpr int kBlockStreamPanel = 16;   // most output blocks one work entry walks
constexpr int kBlockStreamWaves = 65536; // work entries above which one entry walks several blocks
// One wave's work: output blocks order[begin .. begin + count) (they share
// the big operand's open sector tuple, so the rows they read coincide) at
// big-side rows j0 .. j0 + kBlockStreamRows - 1.
struct StreamWork {
  int32_t begin, count, j0, pad;
};
// coef (coupled steps, else null): a real factor per segment.
hipError_t launch_zstream_blocks(bool big_is_a, const void* A, const void* B, void* C,
                                 const GemmItem* items, const GemmSeg* segs, const uint32_t* koff,
                                 const int32_t* order, const StreamWork* work, int nwork, int conjA,
                                 int conjB, hipStream_t stream, const double* coef = nullptr);

// Lean form of the per-block streaming step (the small operand staged in LDS
// when it fits, else read at wave-uniform addresses through the caches),
// used when every output block has at most kLeanMaxS small-side outputs
// and kLeanRank big-side legs, and every segment contracts at most
// kLeanInlineK indices.  Everything a lane needs beyond its row index is precomputed on
// the host: row digits come from a multiply-high by a per-leg reciprocal
// (exact while J * dim < 2^32), the output offsets of the small-side index
// per block, and per segment the small-side offsets inside the small
// operand, which every workgroup stages in LDS once.
constexpr int kLeanMaxS = 16;
constexpr int kLeanSmallMax = 4096;      // small-operand elements staged in LDS (64 KB)
constexpr int kLeanRank = 4;             // big-side legs (unit legs dropped)
constexpr int kLeanInlineK = 8;          // contracted extent of a segment
constexpr int kLeanRows = 2;             // big-side rows per lane (64 * kLeanRows per entry)
// Fixed-width records, read as whole 16-byte vectors (no per-field loads).
struct alignas(16) LeanItem {
  int64_t c_off;
  int32_t J, S, seg_begin, seg_count, rank, pad;
  uint32_t dims[kLeanRank];     // big-side legs
  uint32_t magic[kLeanRank];    // ceil(2^32 / dims[l]) (unused for the last leg)
  uint32_t cstr[kLeanRank];     // output strides of the big-side legs
  uint32_t cso[kLeanMaxS];      // output offset of small-side index s
};
struct alignas(16) LeanSeg {
  int64_t big_off;              // big block start
  double coef;                  // factor of this block pair (1 but in coupled steps)
  int32_t K, pad[3];
  uint32_t bstr[kLeanRank];     // big block strides of the big-side legs
  uint32_t sso[kLeanMaxS];      // small offset (block start included) of small-side index s
  uint32_t kb[kLeanInlineK];    // big and small offsets of every contracted index k
  uint32_t ks[kLeanInlineK];
};
// One wave's work: rows j0 .. j0 + 64 * kLeanRows - 1 of the output blocks items[begin ..
// begin + count) (the records are stored group by group), which share the
// big operand's open sector tuple: same rows, same row digits.
struct LeanWork {
  int32_t begin, count, j0, pad;
};
// S_MAX: largest S over the items, rounded up to 4, 8 or 16.
hipError_t launch_zstream_lean(bool big_is_a, int smax, const void* A, const void* B, void* C,
                               const LeanItem* items, const LeanSeg* segs, const LeanWork* work,
                               int nwork, int nsmall, int conjA, int conjB, hipStream_t stream);

// Block <-> fused-matrix copies of the fused sector path (sector_kernels.hip).
// A PermBlock maps one dense block (contiguous, first leg fastest) of a
// block-sparse tensor onto a strided rectangle of a fused per-charge matrix.
constexpr int kPermChunk = 256;   // elements per chunk of the chunk -> block table
struct PermBlock {
  int64_t base;                   // first element of the block in the blocked tensor
  int64_t size;                   // elements
  int64_t other;                  // fused-matrix offset of the block's first element, -1 = none
  int32_t rank;
  int32_t dims[GRAFT_MAX_RANK];
  int64_t ostr[GRAFT_MAX_RANK];   // fused-matrix stride of every block leg
};
// A rectangle of a fused per-charge matrix that no block of the operand
// covers (a pair of sector tuples without a stored block): it must read as
// zero when a producing step writes that operand straight into the fused
// layout.  Element (i, j) lives at base + i * rstr + j * cstr.
struct HoleRect {
  int64_t base;
  int32_t rows, cols;
  int32_t rstr, cstr;
};
hipError_t launch_zero_holes(const HoleRect* holes, int nholes, void* fused, hipStream_t stream);
// Handoff of a fused step's result into the next fused step's operand
// layout: map[f] = fwd[src[f]] (kMapZero where src[f] is not an element or
// the element has no fused entry).
hipError_t launch_compose_map(const uint32_t* src, int64_t n, const uint32_t* fwd, uint32_t* out,
                              hipStream_t stream);
// Per-element offset maps, built once per plan, and the copies that use them.
// Blocked tensors of the fused path hold fewer than kMapZero elements.
constexpr uint32_t kMapGap = 0xFFFFFFFFu;    // element between blocks: untouched
constexpr uint32_t kMapZero = 0xFFFFFFFEu;   // block without a fused rectangle: zero
hipError_t launch_perm_map(const PermBlock* blocks, int nblocks, const int32_t* chunk_first,
                           int64_t total, uint32_t* map, hipStream_t stream);
// inv[0 .. ninv) = kMapZero, then inv[map[e]] = e for every mapped element.
hipError_t launch_invert_map(const uint32_t* map, int64_t total, uint32_t* inv, int64_t ninv,
                             hipStream_t stream);
// dst[i] = src[map[i]] for i < total (kMapZero: zero, kMapGap: untouched).
hipError_t launch_map_gather(const void* src, void* dst, const uint32_t* map, int64_t total,
                             hipStream_t stream);

// Krylov vector kernels (vector_kernels.hip).
// gate: optional device flag; gate[0] == 0 skips the pass (zero partials, no traffic).
hipError_t launch_dotc_partials(int64_t n, const void* x, const void* y, double* partials,
                                int nblocks, hipStream_t stream, const double* gate = nullptr);
hipError_t launch_reduce_partials(const double* partials, int nblocks, int k, double* out,
                                  hipStream_t stream);
constexpr int kMaxMultiVec = 64;
// Vector pointers passed by value, so launches stay capturable in graphs.
struct PtrList {
  const double2* p[kMaxMultiVec];
};
hipError_t launch_multi_dotc_partials(int64_t n, int k, const PtrList& V, const void* y,
                                      double* partials, int nblocks, hipStream_t stream,
                                      const double* gate = nullptr);
hipError_t launch_axpby(int64_t n, double ar, double ai, const void* x, double br, double bi,
                        void* y, hipStream_t stream);
hipError_t launch_multi_axpy(int64_t n, int k, const PtrList& V, const double* c, double sign,
                             void* y, hipStream_t stream, const double* gate = nullptr);
hipError_t launch_reorth_gate(const double* c, int k, const double* dot2, double eta, double* beta,
                              double* gate, hipStream_t stream);
hipError_t launch_reorth_finish(const double* dot2, const double* gate, double* beta,
                                hipStream_t stream);
hipError_t launch_sqrt_real(const double* dot2, double* out, hipStream_t stream);
constexpr int kMaxCombine = 32;  // outputs of one block combine
hipError_t launch_block_combine(int64_t n, int k, const PtrList& V, int p, const PtrList& O,
                                const double* c, hipStream_t stream);
hipError_t launch_scale_inv(int64_t n, const double* beta, void* x, hipStream_t stream);
constexpr int kReduceBlocks = 1024;

void set_error(const std::string& msg);

}  // namespace graft
