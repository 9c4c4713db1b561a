// Runtime of the effective-Hamiltonian engine: contexts, the device block
// store, plan compilation (pairwise steps -> grouped strided GEMM launches
// with a liveness-coloured workspace arena), execution, the tile autotuner
// and the Krylov vector entry points.  Everything exported here is declared
// in include/graft_heff.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <functional>
#include <map>
#include <unordered_map>
#include <unordered_set>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "heff_internal.h"

namespace graft {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

#define GRAFT_CHECK_HIP(expr)                                                          \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      set_error(std::string("HIP error in ") + #expr + ": " + hipGetErrorString(_e));  \
      return GRAFT_E_HIP;                                                              \
    }                                                                                  \
  } while (0)

#define GRAFT_REQUIRE(cond, msg)  \
  do {                            \
    if (!(cond)) {                \
      set_error(msg);             \
      return GRAFT_E_INVALID;     \
    }                             \
  } while (0)

// ---------------------------------------------------------------------------
// Tensor structure: legs with sector sizes, blocks with sector indices,
// offsets, dims and strides (elements).  Dense = one sector per leg, one block.
// ---------------------------------------------------------------------------
// Fixed-capacity vector of at most GRAFT_MAX_RANK entries: block sector
// tuples, dims and strides without a heap allocation per block.
template <typename T>
struct SmallVec {
  T v[GRAFT_MAX_RANK];
  int32_t n = 0;
  SmallVec() = default;
  explicit SmallVec(size_t k) : n((int32_t)k) {
    for (int32_t i = 0; i < n; ++i) v[i] = T();
  }
  size_t size() const { return (size_t)n; }
  bool empty() const { return n == 0; }
  void resize(size_t k) {
    for (int32_t i = n; i < (int32_t)k; ++i) v[i] = T();
    n = (int32_t)k;
  }
  void assign(size_t k, T x) {
    n = (int32_t)k;
    for (int32_t i = 0; i < n; ++i) v[i] = x;
  }
  void push_back(T x) { v[n++] = x; }
  T& operator[](size_t i) { return v[i]; }
  const T& operator[](size_t i) const { return v[i]; }
  T* begin() { return v; }
  T* end() { return v + n; }
  const T* begin() const { return v; }
  const T* end() const { return v + n; }
  bool operator==(const SmallVec& o) const {
    if (n != o.n) return false;
    for (int32_t i = 0; i < n; ++i)
      if (v[i] != o.v[i]) return false;
    return true;
  }
  bool operator!=(const SmallVec& o) const { return !(*this == o); }
  bool operator<(const SmallVec& o) const {
    const int32_t m = n < o.n ? n : o.n;
    for (int32_t i = 0; i < m; ++i)
      if (v[i] != o.v[i]) return v[i] < o.v[i];
    return n < o.n;
  }
};
using SecKey = SmallVec<int32_t>;

struct Block {
  SecKey sec;
  SmallVec<int64_t> dims;
  SmallVec<int64_t> strides;
  int64_t offset = 0;
  int64_t size = 1;
};

struct SecHash {
  size_t operator()(const SecKey& v) const {
    uint64_t h = 1469598103934665603ull;   // FNV-1a over the sector indices
    for (int32_t x : v) {
      h ^= (uint32_t)x;
      h *= 1099511628211ull;
    }
    return (size_t)h;
  }
};
using SecIndex = std::unordered_map<SecKey, int, SecHash>;

struct TStruct {
  int rank = 0;
  std::vector<std::vector<int64_t>> secdims;  // per leg
  std::vector<Block> blocks;
  SecIndex index;
  int64_t total = 0;  // elements spanned by the buffer
};

// allow_dup: blocks may share a sector tuple (tensors of a coupled plan,
// whose blocks are told apart by fusion-tree labels the descriptor does not
// carry); such structures are addressed by block position, not by index.
static int build_struct(const graft_tensor_desc& d, TStruct& t, std::string& err,
                        bool allow_dup = false) {
  if (d.rank < 0 || d.rank > GRAFT_MAX_RANK) {
    err = "tensor rank out of range";
    return 1;
  }
  t.rank = d.rank;
  t.secdims.assign(d.rank, {});
  const bool dense = (d.sector_dims == nullptr);
  int64_t pos = 0;
  for (int l = 0; l < d.rank; ++l) {
    if (dense) {
      if (d.dims[l] <= 0) {
        err = "non-positive leg dimension";
        return 1;
      }
      t.secdims[l].push_back(d.dims[l]);
    } else {
      if (d.nsectors[l] <= 0) {
        err = "leg without sectors";
        return 1;
      }
      int64_t s = 0;
      for (int q = 0; q < d.nsectors[l]; ++q) {
        t.secdims[l].push_back(d.sector_dims[pos + q]);
        s += d.sector_dims[pos + q];
      }
      pos += d.nsectors[l];
      if (s != d.dims[l]) {
        err = "sector sizes do not sum to the leg dimension";
        return 1;
      }
    }
  }
  const int64_t nb = dense ? 1 : d.nblocks;
  if (nb < 0 || (!dense && nb > 0 && d.block_sectors == nullptr)) {
    err = "sector tensor without block list";
    return 1;
  }
  int64_t next = 0;
  t.total = 0;
  for (int64_t b = 0; b < nb; ++b) {
    Block blk;
    blk.sec.resize(d.rank);
    blk.dims.resize(d.rank);
    blk.strides.resize(d.rank);
    int64_t stride = 1;
    for (int l = 0; l < d.rank; ++l) {
      const int32_t s = dense ? 0 : d.block_sectors[b * d.rank + l];
      if (s < 0 || s >= (int32_t)t.secdims[l].size()) {
        err = "block sector index out of range";
        return 1;
      }
      blk.sec[l] = s;
      blk.dims[l] = t.secdims[l][s];
      if (d.block_strides != nullptr && !dense) {
        blk.strides[l] = d.block_strides[b * d.rank + l];
      } else {
        blk.strides[l] = stride;
      }
      stride *= blk.dims[l];
    }
    blk.size = stride;
    blk.offset = (d.block_offsets != nullptr && !dense) ? d.block_offsets[b] : next;
    next = blk.offset + blk.size;
    int64_t span = 1;
    for (int l = 0; l < d.rank; ++l) span += (blk.dims[l] - 1) * blk.strides[l];
    t.total = std::max(t.total, blk.offset + span);
    if (!allow_dup) {
      if (t.index.count(blk.sec)) {
        err = "duplicate block";
        return 1;
      }
      t.index[blk.sec] = (int)t.blocks.size();
    }
    t.blocks.push_back(std::move(blk));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Compiled step / plan
// ---------------------------------------------------------------------------
struct StepExec {
  int a = 0, b = 0, dst = 0;
  int conjA = 0, conjB = 0;
  int config = 0;
  int pinned = -1;
  std::vector<GemmItem> items;
  std::vector<GemmSeg> segs;
  std::vector<uint32_t> koff;
  int64_t maxM = 0, maxK = 0, maxN = 0;
  double macs = 0.0;
  // device copies
  GemmItem* d_items = nullptr;
  GemmSeg* d_segs = nullptr;
  uint32_t* d_koff = nullptr;
  int total_tiles = 0;
  int ksplit = 1;             // split-K factor (single-block steps with few tiles)
  bool tr = false;            // transposed MFMA orientation in use
  std::vector<int> kfast_memo;   // per tile configuration: piped_kfast, computed on first use
  int xcd_config = -1;           // configuration the XCD split below was computed for
  XcdSplit xcd{};                // work-balanced tile ranges per XCD
  bool tr_auto = false;       // orientation from the output's unit-stride leg
  int tr_pinned = -1;         // -1 = automatic
  int skinny = 0;             // streaming kernel eligible: 1 = A is the large operand, 2 = B
  // Per-block streaming (multi-block steps): 1 = A is the large operand, 2 = B.
  int bstream = 0;
  bool bstream_pref = false;          // the default choice for this step
  std::vector<int32_t> sorder;        // output blocks grouped by the large operand's open tuple
  std::vector<int32_t> spanel;        // group starts in sorder, then sorder.size()
  std::vector<StreamWork> swork;      // one entry per wave
  int32_t* d_sorder = nullptr;
  StreamWork* d_swork = nullptr;
  // Lean form of the per-block streaming step (small operand staged in LDS)
  bool lean = false;
  int lean_smax = 0;                  // largest small-side extent of an output block
  int lean_rank = 0;                  // most big-side legs of extent > 1 of an output block
  const char* lean_why = "";          // why a streaming step has no lean records (plan timing)
  int64_t small_total = 0;            // elements of the small operand
  std::vector<LeanItem> litems;
  std::vector<LeanSeg> lsegs;
  std::vector<LeanWork> lwork;
  LeanItem* d_litems = nullptr;
  LeanSeg* d_lsegs = nullptr;
  LeanWork* d_lwork = nullptr;
  uint64_t partial_bytes = 0; // workspace for the split-K partial tiles
  // Coupled steps with few output tiles: every output block's pair list cut
  // into cslices chunks, chunk s storing into slice s of the partial buffer
  // (ctot elements per slice), the slices summed into C afterwards.
  std::vector<GemmItem> items_base;   // the per-block items before the cut
  int cslices = 1;
  int64_t ctot = 0;
  // Fused sector path: the GEMM items above run on per-charge dense matrices
  // gathered from A and B and scattered into C.
  bool fused = false;
  std::vector<PermBlock> pa, pb, pc;
  std::vector<int32_t> ca, cb, cc;       // chunk -> first candidate block
  int64_t fa = 0, fb = 0, fc = 0;        // fused matrix elements of A, B, C
  int64_t totA = 0, totB = 0, totC = 0;  // element ranges of the blocked tensors
  PermBlock* d_pa = nullptr;
  PermBlock* d_pb = nullptr;
  PermBlock* d_pc = nullptr;
  int32_t* d_ca = nullptr;
  int32_t* d_cb = nullptr;
  int32_t* d_cc = nullptr;
  uint32_t* d_ma = nullptr;              // per-element fused offsets (built once per plan)
  uint32_t* d_mb = nullptr;
  uint32_t* d_mc = nullptr;
  uint32_t* d_mci = nullptr;             // fused C offset -> blocked element (mapped stores)
  bool c_complete = false;               // every blocked C element has a fused entry
  uint64_t fused_bytes = 0;
  // Handoff (FP64 mapped path): this step stores its result through d_mho
  // straight into the fused operand layout of step ho_step (side ho_side:
  // 0 A, 1 B), whose hole rectangles are zeroed first; that step then reads
  // the operand in place, with no gather.
  int ho_step = -1, ho_side = 0;
  uint32_t* d_mho = nullptr;
  bool ho_in_a = false, ho_in_b = false;   // an operand arrives in fused layout
  int ho_from_a = -1, ho_from_b = -1;      // the step that produces it
  std::vector<HoleRect> holes_a, holes_b;  // fused rectangles no operand block covers
  HoleRect* d_holes_a = nullptr;
  HoleRect* d_holes_b = nullptr;
  // Split-word handoff (BF16X3 applies): a streaming step whose result is
  // the A operand of a tiled GEMM step writes it as split-bf16 word pairs
  // (pl_to: that step) and the GEMM stages them without conversion, its B
  // split once per apply into the plane region (pl_from: the producer).
  int pl_to = -1, pl_from = -1;
  uint64_t bplane_bytes = 0;
  // Coupled step (non-abelian blocks): one real factor per segment, applied
  // to the A tile as it is staged; FP64 single-buffered tiles only.
  std::vector<double> coef;
  double* d_coef = nullptr;
};

struct graft_plan_impl {
  int nslots = 0, output_slot = 0, ninputs = 0;
  std::vector<TStruct> slots;
  std::vector<StepExec> steps;
  std::vector<int64_t> ws_offset;  // per slot byte offset in workspace (-1 = not in workspace)
  uint64_t ws_slots = 0;           // bytes of the intermediate slots
  uint64_t ws_bytes = 0;           // slots + split-K partial region + fused matrices
  uint64_t ws_partial = 0;         // byte offset of the split-K partial region
  uint64_t ws_fused = 0;           // byte offset of the fused-matrix region
  uint64_t ws_planes = 0;          // byte offset of the split-word B region
  double macs = 0.0;
  void* d_meta = nullptr;  // one allocation holding every step's items/segs/koff
  void* d_maps = nullptr;  // element -> fused-offset maps of every fused step
  uint64_t meta_bytes = 0, map_bytes = 0;
  int device = 0;
};

}  // namespace graft

struct graft_plan {
  graft::graft_plan_impl impl;
};

struct graft_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* ws = nullptr;
  uint64_t ws_size = 0;
  uint64_t ws_cap = 0;
  double* d_partials = nullptr;  // reduction scratch
  double* d_result = nullptr;
  double* h_result = nullptr;    // pinned
  double* d_coef = nullptr;
  double* d_comb = nullptr;      // block-combine coefficients
  bool capturing = false;        // a graph capture is open on the stream
  int sector_fuse = 1;  // 0 per-block GEMM, 1 automatic, 2 fused / 3 per-block streaming for every multi-block step
  int fused_map_loads = 1;  // fused FP64 MFMA steps read / write the blocks through the maps
  int lean_stream = 1;      // per-block streaming steps use the lean kernel where eligible
  int split_words = 1;      // BF16X3: streaming producers write split-bf16 words for their GEMM
  uint64_t ws_generation = 0;    // bumped whenever the workspace moves
};

struct graft_graph {
  hipGraphExec_t exec = nullptr;
  graft_ctx* ctx = nullptr;
  uint64_t ws_generation = 0;    // the workspace the captured launches address
};

namespace graft {

static int set_device(graft_ctx* ctx) {
  GRAFT_CHECK_HIP(hipSetDevice(ctx->device));
  return GRAFT_OK;
}

// Merge adjacent legs of one index group (m, n or k) when they are
// contiguous in every tensor addressing that group: legs l-1, l merge when
// every stride set (C's and each contribution's) steps by the dimension.
static void merge_legs(SmallVec<int64_t>& dims, SmallVec<int64_t>& cs, SmallVec<int64_t>* ss,
                       size_t nss) {
  int32_t w = 0;
  for (int32_t l = 0; l < dims.n; ++l) {
    const int64_t d = dims[l];
    if (d == 1) continue;
    bool merge = w > 0 && cs[l] == cs[w - 1] * dims[w - 1];
    for (size_t t = 0; merge && t < nss; ++t) merge = ss[t][l] == ss[t][w - 1] * dims[w - 1];
    if (merge) {
      dims[w - 1] *= d;
    } else {
      dims[w] = d;
      cs[w] = cs[l];
      for (size_t t = 0; t < nss; ++t) ss[t][w] = ss[t][l];
      ++w;
    }
  }
  dims.n = w;
  cs.n = w;
  for (size_t t = 0; t < nss; ++t) ss[t].n = w;
}

static int choose_config(int64_t M, int64_t N, int skinny) {
  if (skinny == 1 && M >= 16384) return kSkinnyConfig;
  if (skinny == 2 && N >= 16384) return kSkinnyConfig;
  if (N <= 16) return 3;
  if (M <= 16) return 6;
  if (M <= 32 && N <= 32) return 5;
  // Measured on MI355X (autotuner, root-edge eff_h2 at D = 256): the 64x64
  // three-product tiles beat the 128x128 four-product ones on every
  // MFMA-bound step, and the double-buffered kernel (pipelined on
  // whole-tile steps) beats the single-buffered one in every load direction,
  // so it is the default for large blocks.
  if (M >= 64 && N >= 64) return 15;
  if (N <= 64 && M >= 128) return 1;
  if (M <= 64 && N >= 128) return 4;
  return 2;
}

// Load directions of a step for the pipelined tiles of configuration c:
// bit 0 every segment loads A along k, bit 1 B; -1 unless every tile of c is
// whole (M, N and K multiples of the tile), there is no conjugation and all
// segments share their directions.
static int piped_kfast(StepExec& ex, int c) {
  if (c < 0 || c >= kNumTileConfigs || ex.conjA || ex.conjB || ex.ksplit != 1) return -1;
  if (ex.kfast_memo.empty()) ex.kfast_memo.assign(kNumTileConfigs, -2);
  if (ex.kfast_memo[c] != -2) return ex.kfast_memo[c];
  const TileConfig& tc = kTileConfigs[c];
  int kf = -1;
  bool ok = true;
  for (const GemmItem& it : ex.items) {
    if (it.M % tc.bm || it.N % tc.bn) { ok = false; break; }
    for (int s = it.seg_begin; ok && s < it.seg_begin + it.seg_count; ++s) {
      const GemmSeg& sg = ex.segs[s];
      const int f = (sg.a_kfast ? 1 : 0) | (sg.b_kfast ? 2 : 0);
      if (sg.K % tc.bk || (kf >= 0 && f != kf)) ok = false;
      kf = f;
    }
    if (!ok) break;
  }
  ex.kfast_memo[c] = ok ? kf : -1;
  return ex.kfast_memo[c];
}

// Tile ranges of equal work per XCD for the current tile configuration:
// a tile of an item costs its k-tiles plus a fixed 2 (start-up, epilogue).
// Steps with one item keep equal tile counts.
static const XcdSplit& xcd_split(StepExec& ex) {
  if (ex.xcd_config == ex.config) return ex.xcd;
  ex.xcd_config = ex.config;
  ex.xcd = uniform_xcd_split(ex.total_tiles);
  if (ex.items.size() < 2 || ex.config < 0 || ex.config >= kNumTileConfigs) return ex.xcd;
  const int bk = kTileConfigs[ex.config].bk;
  std::vector<double> w(ex.items.size());
  double total = 0.0;
  for (size_t i = 0; i < ex.items.size(); ++i) {
    const GemmItem& it = ex.items[i];
    int64_t kt = 0;
    for (int s = it.seg_begin; s < it.seg_begin + it.seg_count; ++s) kt += (ex.segs[s].K + bk - 1) / bk;
    w[i] = (double)(kt + 2);
    const int next = i + 1 < ex.items.size() ? ex.items[i + 1].tile_begin : ex.total_tiles;
    total += w[i] * (next - it.tile_begin);
  }
  double cum = 0.0;
  size_t i = 0;
  for (int x = 1; x < 8; ++x) {
    const double target = total * x / 8.0;
    while (i < ex.items.size()) {
      const int next = i + 1 < ex.items.size() ? ex.items[i + 1].tile_begin : ex.total_tiles;
      const double iw = w[i] * (next - ex.items[i].tile_begin);
      if (cum + iw >= target) break;
      cum += iw;
      ++i;
    }
    int t = ex.total_tiles;
    if (i < ex.items.size()) {
      const int next = i + 1 < ex.items.size() ? ex.items[i + 1].tile_begin : ex.total_tiles;
      t = ex.items[i].tile_begin + (int)std::ceil((target - cum) / w[i]);
      t = std::min(t, next);
    }
    ex.xcd.start[x] = std::max(t, ex.xcd.start[x - 1]);
  }
  ex.xcd.start[0] = 0;
  ex.xcd.start[8] = ex.total_tiles;
  return ex.xcd;
}

static bool canonical_blocks(const TStruct& T) {
  for (const auto& b : T.blocks) {
    if (b.size >= (1ll << 31)) return false;   // the block copies index blocks in 32 bits
    int64_t s = 1;
    for (size_t l = 0; l < b.dims.size(); ++l) {
      if (b.dims[l] > 1 && b.strides[l] != s) return false;
      s *= b.dims[l];
    }
  }
  return true;
}

static std::vector<int32_t> chunk_table(const std::vector<PermBlock>& blocks, int64_t total) {
  const int64_t nchunks = (total + kPermChunk - 1) / kPermChunk;
  std::vector<int32_t> t(nchunks + 1, 0);
  size_t b = 0;
  for (int64_t c = 0; c <= nchunks; ++c) {
    const int64_t e = std::min<int64_t>(c * kPermChunk, std::max<int64_t>(total - 1, 0));
    while (b + 1 < blocks.size() && blocks[b + 1].base <= e) ++b;
    t[c] = (int32_t)b;
  }
  // a chunk's candidates run up to the block that owns its last element
  for (int64_t c = 0; c < nchunks; ++c) {
    const int64_t last = std::min<int64_t>((c + 1) * kPermChunk - 1, total - 1);
    size_t hb = t[c];
    while (hb + 1 < blocks.size() && blocks[hb + 1].base <= last) ++hb;
    t[c + 1] = std::max<int32_t>(t[c + 1], (int32_t)hb);
  }
  return t;
}

// Per-charge fusion of one step (TensorKit's matrix blocks by coupled sector):
// union-find over A-open tuples r, contracted tuples c and B-open tuples s,
// linked by the stored A blocks (r, c) and B blocks (c, s).  Every component
// with both A and B blocks becomes one dense GEMM C_g = A_g B_g with rows
// (r, open index), contraction (c, contracted index), columns (s, open index).
static int build_fused(const graft_pair_step& st, const TStruct& A, const TStruct& B,
                       const TStruct& C, const std::vector<int>& pos_of_out, StepExec& ex,
                       std::string& err, bool want_holes_a, bool want_holes_b) {
  const int nOA = st.n_oindA, nCA = st.n_cindA, nOB = st.n_oindB;
  using Key = SecKey;
  // Tuple ids in order of first appearance over the blocks (deterministic).
  struct Ids {
    std::unordered_map<Key, int, SecHash> id;
    std::vector<Key> keys;
    int get(const Key& k) {
      auto f = id.find(k);
      if (f != id.end()) return f->second;
      const int v = (int)keys.size();
      id.emplace(k, v);
      keys.push_back(k);
      return v;
    }
    int find(const Key& k) const {
      auto f = id.find(k);
      return f == id.end() ? -1 : f->second;
    }
  };
  Ids R, Cc, S;
  std::vector<int> aR(A.blocks.size()), aC(A.blocks.size()), bC(B.blocks.size()), bS(B.blocks.size());
  Key key;
  for (size_t i = 0; i < A.blocks.size(); ++i) {
    key.assign(nOA, 0);
    for (int p = 0; p < nOA; ++p) key[p] = A.blocks[i].sec[st.oindA[p]];
    aR[i] = R.get(key);
    key.assign(nCA, 0);
    for (int q = 0; q < nCA; ++q) key[q] = A.blocks[i].sec[st.cindA[q]];
    aC[i] = Cc.get(key);
  }
  for (size_t i = 0; i < B.blocks.size(); ++i) {
    key.assign(nCA, 0);
    for (int q = 0; q < nCA; ++q) key[q] = B.blocks[i].sec[st.cindB[q]];
    bC[i] = Cc.get(key);
    key.assign(nOB, 0);
    for (int p = 0; p < nOB; ++p) key[p] = B.blocks[i].sec[st.oindB[p]];
    bS[i] = S.get(key);
  }
  const int nr = (int)R.keys.size(), nc = (int)Cc.keys.size(), ns = (int)S.keys.size();
  std::vector<int> par(nr + nc + ns);
  std::iota(par.begin(), par.end(), 0);
  auto find = [&](int x) {
    while (par[x] != x) {
      par[x] = par[par[x]];
      x = par[x];
    }
    return x;
  };
  auto unite = [&](int x, int y) { par[find(x)] = find(y); };
  for (size_t i = 0; i < A.blocks.size(); ++i) unite(aR[i], nr + aC[i]);
  for (size_t i = 0; i < B.blocks.size(); ++i) unite(nr + bC[i], nr + nc + bS[i]);
  std::vector<char> hasA(nr + nc + ns, 0), hasB(nr + nc + ns, 0);
  for (size_t i = 0; i < A.blocks.size(); ++i) hasA[find(aR[i])] = 1;
  for (size_t i = 0; i < B.blocks.size(); ++i) hasB[find(nr + bC[i])] = 1;
  auto tuple_size = [](const TStruct& T, const int32_t* legs, const Key& k) {
    int64_t n = 1;
    for (size_t i = 0; i < k.size(); ++i) n *= T.secdims[legs[i]][k[i]];
    return n;
  };
  struct Comp {
    int64_t M = 0, K = 0, N = 0, aoff = 0, boff = 0, coff = 0;
  };
  std::vector<Comp> comps;
  std::unordered_map<int, int> comp;   // union-find root -> component index
  auto comp_of = [&](int root) {
    auto f = comp.find(root);
    if (f != comp.end()) return f->second;
    const int g = (int)comps.size();
    comp.emplace(root, g);
    comps.emplace_back();
    return g;
  };
  // (component, offset) of every tuple id; component -1 = no GEMM
  std::vector<int> rg(nr, -1), cg(nc, -1), sg(ns, -1);
  std::vector<int64_t> ro(nr, 0), co(nc, 0), so(ns, 0);
  for (int r = 0; r < nr; ++r) {
    const int root = find(r);
    if (!hasA[root] || !hasB[root]) continue;
    const int gi = comp_of(root);
    rg[r] = gi;
    ro[r] = comps[gi].M;
    comps[gi].M += tuple_size(A, st.oindA, R.keys[r]);
  }
  for (int c = 0; c < nc; ++c) {
    const int root = find(nr + c);
    if (!hasA[root] || !hasB[root]) continue;
    const int gi = comp_of(root);
    cg[c] = gi;
    co[c] = comps[gi].K;
    comps[gi].K += tuple_size(A, st.cindA, Cc.keys[c]);
  }
  for (int q = 0; q < ns; ++q) {
    const int root = find(nr + nc + q);
    if (!hasA[root] || !hasB[root]) continue;
    const int gi = comp_of(root);
    sg[q] = gi;
    so[q] = comps[gi].N;
    comps[gi].N += tuple_size(B, st.oindB, S.keys[q]);
  }
  int64_t fa = 0, fb = 0, fc = 0;
  for (auto& g : comps) {
    g.aoff = fa;
    g.boff = fb;
    g.coff = fc;
    fa += g.M * g.K;
    fb += g.K * g.N;
    fc += g.M * g.N;
  }
  if (comps.empty()) {
    err = "contraction step has no matching blocks";
    return 1;
  }
  if (fa >= (1ll << 32) - 2 || fb >= (1ll << 32) - 2 || fc >= (1ll << 32) - 2) {
    err = "fused sector matrices exceed 2^32 elements";
    return 1;
  }
  // Mixed-radix strides of the legs of a tuple, the legs taken in `order`
  // (positions into `legs`).  Rows, contracted indices and columns follow the
  // memory order of the operand that owns them, so a block's consecutive
  // elements tend to land on consecutive fused-matrix entries.
  auto order_of = [](const int32_t* legs, int n) {
    std::vector<int> o(n);
    std::iota(o.begin(), o.end(), 0);
    std::sort(o.begin(), o.end(), [&](int x, int y) { return legs[x] < legs[y]; });
    return o;
  };
  const std::vector<int> rorder = order_of(st.oindA, nOA);
  const std::vector<int> sorder = order_of(st.oindB, nOB);
  const std::vector<int> korder = A.total >= B.total ? order_of(st.cindA, nCA) : order_of(st.cindB, nCA);
  auto radix_o = [](const TStruct& T, const int32_t* legs, const Key& k, const std::vector<int>& ord,
                    int64_t* r) {
    int64_t s = 1;
    for (int i : ord) {
      r[i] = s;
      s *= T.secdims[legs[i]][k[i]];
    }
  };
  auto sorted_blocks = [](const TStruct& T) {
    std::vector<int> idx(T.blocks.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return T.blocks[x].offset < T.blocks[y].offset; });
    return idx;
  };
  auto make_pb = [](const Block& b) {
    PermBlock pbk;
    std::memset(&pbk, 0, sizeof(pbk));
    pbk.base = b.offset;
    pbk.size = b.size;
    pbk.other = -1;
    pbk.rank = (int32_t)b.dims.size();
    for (size_t l = 0; l < b.dims.size(); ++l) pbk.dims[l] = (int32_t)b.dims[l];
    return pbk;
  };
  // Fused-matrix orientation per operand, chosen to minimize the number of
  // separate runs the block copies write or read: a run continues while the
  // next block leg (in memory order, unit legs skipped) steps the fused offset
  // by the run length so far.
  auto runs = [](const Block& b, const int64_t* ostr) {
    int64_t run = 1;
    for (size_t l = 0; l < b.dims.size(); ++l) {
      if (b.dims[l] == 1) continue;
      if (ostr[l] != run) break;
      run *= b.dims[l];
    }
    return (double)b.size / (double)run;
  };
  int64_t rr[GRAFT_MAX_RANK], rk[GRAFT_MAX_RANK], rs[GRAFT_MAX_RANK];
  int64_t o1[GRAFT_MAX_RANK], o2[GRAFT_MAX_RANK];
  // A: (m, k) at m*K + k (a_kf) or m + M*k
  auto a_strides = [&](int i, bool kf, int64_t* o) {
    const Comp& g = comps[rg[aR[i]]];
    radix_o(A, st.oindA, R.keys[aR[i]], rorder, rr);
    radix_o(A, st.cindA, Cc.keys[aC[i]], korder, rk);
    for (int p = 0; p < nOA; ++p) o[st.oindA[p]] = rr[p] * (kf ? g.K : 1);
    for (int q = 0; q < nCA; ++q) o[st.cindA[q]] = rk[q] * (kf ? 1 : g.M);
  };
  // B: (k, n) at k + K*n (b_kf) or n + N*k
  auto b_strides = [&](int i, bool kf, int64_t* o) {
    const Comp& g = comps[cg[bC[i]]];
    radix_o(B, st.cindB, Cc.keys[bC[i]], korder, rk);
    radix_o(B, st.oindB, S.keys[bS[i]], sorder, rs);
    for (int q = 0; q < nCA; ++q) o[st.cindB[q]] = rk[q] * (kf ? 1 : g.N);
    for (int p = 0; p < nOB; ++p) o[st.oindB[p]] = rs[p] * (kf ? g.K : 1);
  };
  // C: (m, n) at m + M*n (c_mf) or n + N*m
  Key rkey(nOA), skey(nOB);
  auto c_lookup = [&](const Block& b, int& r, int& q) {
    for (int p = 0; p < nOA; ++p) rkey[p] = b.sec[pos_of_out[p]];
    for (int p = 0; p < nOB; ++p) skey[p] = b.sec[pos_of_out[nOA + p]];
    r = R.find(rkey);
    q = S.find(skey);
    return r >= 0 && q >= 0 && rg[r] >= 0 && rg[r] == sg[q];
  };
  // (reads the tuple keys c_lookup just filled for this block)
  auto c_strides = [&](const Block& b, int r, bool mf, int64_t* o) {
    const Comp& g = comps[rg[r]];
    radix_o(A, st.oindA, rkey, rorder, rr);
    radix_o(B, st.oindB, skey, sorder, rs);
    for (int qq = 0; qq < (int)b.dims.size(); ++qq) {
      const int p = st.out[qq];
      o[qq] = p < nOA ? rr[p] * (mf ? 1 : g.N) : rs[p - nOA] * (mf ? g.M : 1);
    }
  };
  double ra1 = 0, ra2 = 0, rb1 = 0, rb2 = 0, rc1 = 0, rc2 = 0;
  for (size_t i = 0; i < A.blocks.size(); ++i) {
    if (rg[aR[i]] < 0 || rg[aR[i]] != cg[aC[i]]) continue;
    a_strides((int)i, false, o1);
    a_strides((int)i, true, o2);
    ra1 += runs(A.blocks[i], o1);
    ra2 += runs(A.blocks[i], o2);
  }
  for (size_t i = 0; i < B.blocks.size(); ++i) {
    if (cg[bC[i]] < 0 || cg[bC[i]] != sg[bS[i]]) continue;
    b_strides((int)i, false, o1);
    b_strides((int)i, true, o2);
    rb1 += runs(B.blocks[i], o1);
    rb2 += runs(B.blocks[i], o2);
  }
  for (size_t i = 0; i < C.blocks.size(); ++i) {
    int r, q;
    if (!c_lookup(C.blocks[i], r, q)) continue;
    c_strides(C.blocks[i], r, true, o1);
    c_strides(C.blocks[i], r, false, o2);
    rc1 += runs(C.blocks[i], o1);
    rc2 += runs(C.blocks[i], o2);
  }
  const bool a_kf = ra2 < ra1;
  const bool b_kf = rb2 < rb1;
  const bool c_mf = rc1 <= rc2;
  ex.pa.clear();
  ex.pb.clear();
  ex.pc.clear();
  ex.pa.reserve(A.blocks.size());
  ex.pb.reserve(B.blocks.size());
  ex.pc.reserve(C.blocks.size());
  for (int i : sorted_blocks(A)) {
    const Block& b = A.blocks[i];
    PermBlock pbk = make_pb(b);
    const int g1 = rg[aR[i]], g2 = cg[aC[i]];
    if (g1 >= 0 && g1 == g2) {
      const Comp& g = comps[g1];
      pbk.other = g.aoff + ro[aR[i]] * (a_kf ? g.K : 1) + co[aC[i]] * (a_kf ? 1 : g.M);
      a_strides(i, a_kf, o1);
      for (size_t l = 0; l < b.dims.size(); ++l) pbk.ostr[l] = o1[l];
    }
    ex.pa.push_back(pbk);
  }
  for (int i : sorted_blocks(B)) {
    const Block& b = B.blocks[i];
    PermBlock pbk = make_pb(b);
    const int g1 = cg[bC[i]], g2 = sg[bS[i]];
    if (g1 >= 0 && g1 == g2) {
      const Comp& g = comps[g1];
      pbk.other = g.boff + co[bC[i]] * (b_kf ? 1 : g.N) + so[bS[i]] * (b_kf ? g.K : 1);
      b_strides(i, b_kf, o1);
      for (size_t l = 0; l < b.dims.size(); ++l) pbk.ostr[l] = o1[l];
    }
    ex.pb.push_back(pbk);
  }
  for (int i : sorted_blocks(C)) {
    const Block& b = C.blocks[i];
    PermBlock pbk = make_pb(b);
    int r, q;
    if (c_lookup(b, r, q)) {
      const Comp& g = comps[rg[r]];
      pbk.other = g.coff + ro[r] * (c_mf ? 1 : g.N) + so[q] * (c_mf ? g.M : 1);
      c_strides(b, r, c_mf, o1);
      for (size_t l = 0; l < b.dims.size(); ++l) pbk.ostr[l] = o1[l];
    }
    ex.pc.push_back(pbk);
  }
  ex.c_complete = std::all_of(ex.pc.begin(), ex.pc.end(),
                              [](const PermBlock& b) { return b.other >= 0; });
  // Hole rectangles of the fused operands: tuple pairs of a component with no
  // stored block (only for operands another step may produce).
  auto holes = [&](bool side_a, std::vector<HoleRect>& out) {
    out.clear();
    const int n1 = side_a ? nr : nc, n2 = side_a ? nc : ns;
    // position of every tuple id inside its component's tuple list, and a
    // presence bitmap per component (tuple pairs with a stored block)
    std::vector<std::vector<int>> t1(comps.size()), t2(comps.size());
    std::vector<int> p1(n1, -1), p2(n2, -1);
    for (int x = 0; x < n1; ++x) {
      const int gi = side_a ? rg[x] : cg[x];
      if (gi >= 0) {
        p1[x] = (int)t1[gi].size();
        t1[gi].push_back(x);
      }
    }
    for (int y = 0; y < n2; ++y) {
      const int gi = side_a ? cg[y] : sg[y];
      if (gi >= 0) {
        p2[y] = (int)t2[gi].size();
        t2[gi].push_back(y);
      }
    }
    std::vector<std::vector<char>> have(comps.size());
    for (size_t gi = 0; gi < comps.size(); ++gi) have[gi].assign(t1[gi].size() * t2[gi].size(), 0);
    const size_t nb = side_a ? A.blocks.size() : B.blocks.size();
    for (size_t i = 0; i < nb; ++i) {
      const int x = side_a ? aR[i] : bC[i], y = side_a ? aC[i] : bS[i];
      const int gi = side_a ? rg[x] : cg[x];
      if (gi < 0 || p1[x] < 0 || p2[y] < 0) continue;
      have[gi][(size_t)p1[x] * t2[gi].size() + p2[y]] = 1;
    }
    for (size_t gi = 0; gi < comps.size(); ++gi) {
      const Comp& g = comps[gi];
      const size_t w = t2[gi].size();
      for (size_t ix = 0; ix < t1[gi].size(); ++ix)
        for (size_t iy = 0; iy < w; ++iy) {
          if (have[gi][ix * w + iy]) continue;
          const int x = t1[gi][ix], y = t2[gi][iy];
          HoleRect h;
          if (side_a) {   // rows r, columns c of A_g: (m, k) at m*K + k or m + M*k
            h.rows = (int32_t)tuple_size(A, st.oindA, R.keys[x]);
            h.cols = (int32_t)tuple_size(A, st.cindA, Cc.keys[y]);
            h.rstr = a_kf ? (int32_t)g.K : 1;
            h.cstr = a_kf ? 1 : (int32_t)g.M;
            h.base = g.aoff + ro[x] * h.rstr + co[y] * h.cstr;
          } else {        // rows c, columns s of B_g: (k, n) at k + K*n or n + N*k
            h.rows = (int32_t)tuple_size(A, st.cindA, Cc.keys[x]);
            h.cols = (int32_t)tuple_size(B, st.oindB, S.keys[y]);
            h.rstr = b_kf ? 1 : (int32_t)g.N;
            h.cstr = b_kf ? (int32_t)g.K : 1;
            h.base = g.boff + co[x] * h.rstr + so[y] * h.cstr;
          }
          out.push_back(h);
        }
    }
  };
  if (want_holes_a) holes(true, ex.holes_a);
  if (want_holes_b) holes(false, ex.holes_b);
  ex.fa = fa;
  ex.fb = fb;
  ex.fc = fc;
  ex.totA = A.total;
  ex.totB = B.total;
  ex.totC = C.total;
  ex.ca = chunk_table(ex.pa, A.total);
  ex.cb = chunk_table(ex.pb, B.total);
  ex.cc = chunk_table(ex.pc, C.total);
  // grouped dense GEMM over the components (column-major fused matrices)
  ex.items.clear();
  ex.segs.clear();
  ex.koff.clear();
  ex.maxM = ex.maxN = ex.maxK = 0;
  for (auto& g : comps) {
    if (g.M == 0 || g.N == 0 || g.K == 0) continue;
    GemmItem it;
    std::memset(&it, 0, sizeof(it));
    it.c_off = g.coff;
    it.M = (int32_t)g.M;
    it.N = (int32_t)g.N;
    it.mrank = 1;
    it.nrank = 1;
    it.mdims[0] = (int32_t)g.M;
    it.ndims[0] = (int32_t)g.N;
    it.c_mstr[0] = c_mf ? 1u : (uint32_t)g.N;
    it.c_nstr[0] = c_mf ? (uint32_t)g.M : 1u;
    it.seg_begin = (int)ex.segs.size();
    it.seg_count = 1;
    GemmSeg sg;
    std::memset(&sg, 0, sizeof(sg));
    sg.a_off = g.aoff;
    sg.b_off = g.boff;
    sg.K = (int32_t)g.K;
    sg.a_mstr[0] = a_kf ? (uint32_t)g.K : 1u;
    sg.b_nstr[0] = b_kf ? (uint32_t)g.K : 1u;
    sg.koff_idx = (int32_t)ex.koff.size();
    for (int64_t k = 0; k < g.K; ++k) ex.koff.push_back((uint32_t)(a_kf ? k : k * g.M));
    for (int64_t k = 0; k < g.K; ++k) ex.koff.push_back((uint32_t)(b_kf ? k : k * g.N));
    sg.a_kfast = a_kf ? 1 : 0;
    sg.b_kfast = b_kf ? 1 : 0;
    ex.segs.push_back(sg);
    ex.items.push_back(it);
    ex.maxM = std::max(ex.maxM, g.M);
    ex.maxN = std::max(ex.maxN, g.N);
    ex.maxK = std::max(ex.maxK, g.K);
  }
  ex.fused = true;
  ex.tr_auto = ex.tr = c_mf;      // transposed MFMA issue when the output is m-fastest
  // Streaming kernel when every component has a short K and one small side
  // on the same operand (the TTNO-absorb steps of a sector map).
  ex.skinny = 0;
  {
    auto fits = [](int64_t K, int64_t S) {
      return (K <= kSkinnyMaxK && S <= kSkinnyMaxS) || (K <= kSkinnyAccMaxK && S <= kSkinnyAccMaxS);
    };
    bool all1 = !ex.items.empty(), all2 = !ex.items.empty();
    for (size_t i = 0; i < ex.items.size(); ++i) {
      const int64_t K = ex.segs[ex.items[i].seg_begin].K, M = ex.items[i].M, N = ex.items[i].N;
      all1 = all1 && fits(K, N) && M > N;
      all2 = all2 && fits(K, M) && N > M;
    }
    ex.skinny = all1 ? 1 : (all2 ? 2 : 0);
  }
  auto al = [](int64_t x) { return (uint64_t)((x * 16 + 255) / 256) * 256; };
  ex.fused_bytes = al(fa) + al(fb) + al(fc);
  return 0;
}

// Lean per-block streaming records (see LeanItem): built when every output
// block has at most kLeanMaxS small-side outputs and at most kLeanRank
// big-side legs of extent > 1, every segment
// contracts at most kLeanInlineK indices, and every big-side digit decodes
// exactly by a multiply-high (J * dim < 2^32).  Records are stored group by
// group (sorder); work entries run rows outer, the blocks of one big-side
// tuple inner.
static constexpr int64_t kLeanWalkAbove = 1 << 21;
static void build_lean(StepExec& ex) {
  ex.lean = false;
  ex.litems.clear();
  ex.lsegs.clear();
  ex.lwork.clear();
  auto fail = [&ex](const char* why) {
    ex.litems.clear();
    ex.lsegs.clear();
    ex.lwork.clear();
    ex.lean_why = why;
  };
  ex.lean_why = "";
  if (ex.bstream == 0 || ex.items.empty()) return;
  const bool bigA = ex.bstream == 1;
  int smax = 1;
  for (const GemmItem& it : ex.items) {
    const int S = bigA ? it.N : it.M;
    if (S > kLeanMaxS) return fail("small side > 16");
    smax = std::max(smax, S);
  }
  for (const GemmSeg& sg : ex.segs)
    if (sg.K > kLeanInlineK) return fail("K > 8");
  ex.litems.reserve(ex.items.size());
  ex.lsegs.reserve(ex.segs.size());
  int lrank = 0;
  for (int32_t item_id : ex.sorder) {
    const GemmItem& it = ex.items[item_id];
    LeanItem li;
    std::memset(&li, 0, sizeof(li));
    li.c_off = it.c_off;
    const int J = bigA ? it.M : it.N;
    const int S = bigA ? it.N : it.M;
    const int brank = bigA ? it.mrank : it.nrank;
    const int srank = bigA ? it.nrank : it.mrank;
    const int32_t* bdims = bigA ? it.mdims : it.ndims;
    const int32_t* sdims = bigA ? it.ndims : it.mdims;
    const uint32_t* cb = bigA ? it.c_mstr : it.c_nstr;
    const uint32_t* cs = bigA ? it.c_nstr : it.c_mstr;
    li.J = J;
    li.S = S;
    // big-side legs of extent > 1 (positions in the item)
    int keep[GL], nk = 0;
    for (int l = 0; l < brank; ++l) {
      if (bdims[l] == 1) continue;
      if ((uint64_t)J * (uint64_t)bdims[l] >= (1ull << 32) || nk == kLeanRank)
        return fail("big-side legs");
      keep[nk++] = l;
    }
    li.rank = nk;
    lrank = std::max(lrank, nk);
    for (int q = 0; q < nk; ++q) {
      const uint32_t d = (uint32_t)bdims[keep[q]];
      li.dims[q] = d;
      li.magic[q] = q + 1 < nk ? (uint32_t)(((1ull << 32) + d - 1) / d) : 0u;
      li.cstr[q] = cb[keep[q]];
    }
    // small-side index s -> digits over the small legs
    int64_t sdig[kLeanMaxS][GL];
    for (int s0 = 0; s0 < S; ++s0) {
      int64_t idx = s0;
      for (int l = 0; l < srank; ++l) {
        sdig[s0][l] = idx % sdims[l];
        idx /= sdims[l];
      }
    }
    for (int s0 = 0; s0 < S; ++s0) {
      uint64_t o = 0;
      for (int l = 0; l < srank; ++l) o += (uint64_t)sdig[s0][l] * cs[l];
      li.cso[s0] = (uint32_t)o;
    }
    li.seg_begin = (int32_t)ex.lsegs.size();
    li.seg_count = it.seg_count;
    for (int g = it.seg_begin; g < it.seg_begin + it.seg_count; ++g) {
      const GemmSeg& sg = ex.segs[g];
      LeanSeg ls;
      std::memset(&ls, 0, sizeof(ls));
      ls.big_off = bigA ? sg.a_off : sg.b_off;
      const int64_t soff = bigA ? sg.b_off : sg.a_off;
      const uint32_t* bstr = bigA ? sg.a_mstr : sg.b_nstr;
      const uint32_t* sstr = bigA ? sg.b_nstr : sg.a_mstr;
      for (int q = 0; q < nk; ++q) ls.bstr[q] = bstr[keep[q]];
      for (int s0 = 0; s0 < S; ++s0) {
        int64_t o = soff;
        for (int l = 0; l < srank; ++l) o += sdig[s0][l] * (int64_t)sstr[l];
        ls.sso[s0] = (uint32_t)o;
      }
      ls.K = sg.K;
      ls.coef = ex.coef.empty() ? 1.0 : ex.coef[g];
      const uint32_t* kA = ex.koff.data() + sg.koff_idx;
      const uint32_t* kbig = bigA ? kA : kA + sg.K;
      const uint32_t* ksml = bigA ? kA + sg.K : kA;
      for (int q = 0; q < sg.K; ++q) {
        ls.kb[q] = kbig[q];
        ls.ks[q] = ksml[q];
      }
      ex.lsegs.push_back(ls);
    }
    ex.litems.push_back(li);
  }
  // Work entries share one row decode, so the blocks of an entry must have
  // the big-side legs of its first block: a group whose blocks merged their
  // legs differently (the merge follows each block's own contributions) is
  // cut into runs of equal legs.
  std::vector<int32_t> lpanel;
  for (size_t g = 0; g + 1 < ex.spanel.size(); ++g) {
    const int b = ex.spanel[g], e = ex.spanel[g + 1];
    int f = b;
    lpanel.push_back(b);
    for (int c = b + 1; c < e; ++c) {
      const LeanItem& x = ex.litems[f];
      const LeanItem& o = ex.litems[c];
      bool same = o.J == x.J && o.rank == x.rank;
      for (int q = 0; same && q < x.rank; ++q) same = o.dims[q] == x.dims[q];
      if (!same) {
        lpanel.push_back(c);
        f = c;
      }
    }
  }
  lpanel.push_back((int32_t)ex.litems.size());
  // One block per entry unless that makes more than kLeanWalkAbove entries:
  // walking the blocks of a run one after another trades memory-level
  // parallelism for cache reuse, which pays only on very large steps.
  int64_t single = 0;
  for (size_t g = 0; g + 1 < lpanel.size(); ++g)
    single += (int64_t)((ex.litems[lpanel[g]].J + 64 * kLeanRows - 1) / (64 * kLeanRows)) *
              (lpanel[g + 1] - lpanel[g]);
  const int panel = (int)std::max<int64_t>(1, (single + kLeanWalkAbove - 1) / kLeanWalkAbove);
  for (size_t g = 0; g + 1 < lpanel.size(); ++g) {
    const int b = lpanel[g], e = lpanel[g + 1];
    const int J = ex.litems[b].J;
    for (int j0 = 0; j0 < J; j0 += 64 * kLeanRows)
      for (int c = b; c < e; c += panel) ex.lwork.push_back({c, std::min(panel, e - c), j0, 0});
  }
  if (ex.lwork.size() >= (1u << 30)) return fail("work entries");
  ex.lean_smax = smax;
  ex.lean_rank = lrank;
  ex.lean = true;
}

// One GEMM item per output block and one segment per contributing block
// pair (the pairs of output block b are pairs[pstart[b] .. pstart[b + 1])).
// Merged index groups use fixed-capacity records: no heap allocation per
// block or pair.
static int build_items(const graft_pair_step& st, const TStruct& A, const TStruct& B,
                       const TStruct& C, const std::vector<int>& pos_of_out,
                       const std::vector<std::pair<int, int>>& pairs,
                       const std::vector<int64_t>& pstart, StepExec& ex, std::string& err) {
  const int nOA = st.n_oindA, nCA = st.n_cindA, nOB = st.n_oindB;
  size_t koff_total = 0;
  for (const auto& pr : pairs) {
    const Block& ab = A.blocks[pr.first];
    int64_t K = 1;
    for (int i = 0; i < nCA; ++i) K *= ab.dims[st.cindA[i]];
    koff_total += 2 * (size_t)K;
  }
  ex.items.reserve(C.blocks.size());
  ex.segs.reserve(pairs.size());
  ex.koff.reserve(koff_total);
  std::vector<SmallVec<int64_t>> am, bn;
  for (int cb = 0; cb < (int)C.blocks.size(); ++cb) {
    const Block& cblk = C.blocks[cb];
    GemmItem it;
    std::memset(&it, 0, sizeof(it));
    it.c_off = cblk.offset;
    it.seg_begin = (int)ex.segs.size();
    const std::pair<int, int>* prs = pairs.data() + pstart[cb];
    const size_t np = (size_t)(pstart[cb + 1] - pstart[cb]);
    // m/n dims are fixed by the output block; strides per contribution.
    SmallVec<int64_t> md, nd, cm0, cn0;
    for (int p = 0; p < nOA; ++p) {
      const int q = pos_of_out[p];
      md.push_back(cblk.dims[q]);
      cm0.push_back(cblk.strides[q]);
    }
    for (int p = 0; p < nOB; ++p) {
      const int q = pos_of_out[nOA + p];
      nd.push_back(cblk.dims[q]);
      cn0.push_back(cblk.strides[q]);
    }
    int64_t M = 1, N = 1;
    for (auto d : md) M *= d;
    for (auto d : nd) N *= d;
    // Merge decisions must hold for every contribution: one common merged
    // layout over C, every A block and every B block of this output block.
    if (am.size() < np) {
      am.resize(np);
      bn.resize(np);
    }
    for (size_t sgi = 0; sgi < np; ++sgi) {
      const Block& ab = A.blocks[prs[sgi].first];
      const Block& bb = B.blocks[prs[sgi].second];
      am[sgi].n = 0;
      bn[sgi].n = 0;
      for (int p = 0; p < nOA; ++p) am[sgi].push_back(ab.strides[st.oindA[p]]);
      for (int p = 0; p < nOB; ++p) bn[sgi].push_back(bb.strides[st.oindB[p]]);
    }
    merge_legs(md, cm0, am.data(), np);
    merge_legs(nd, cn0, bn.data(), np);
    if ((int)md.size() > GL || (int)nd.size() > GL) {
      err = "too many unmerged legs in one GEMM index group";
      return 1;
    }
    it.M = (int32_t)M;
    it.N = (int32_t)N;
    if (M >= (1ll << 31) || N >= (1ll << 31)) {
      err = "GEMM extent exceeds 2^31";
      return 1;
    }
    it.mrank = (int)md.size();
    it.nrank = (int)nd.size();
    for (int l = 0; l < it.mrank; ++l) {
      it.mdims[l] = (int32_t)md[l];
      it.c_mstr[l] = (uint32_t)cm0[l];
    }
    for (int l = 0; l < it.nrank; ++l) {
      it.ndims[l] = (int32_t)nd[l];
      it.c_nstr[l] = (uint32_t)cn0[l];
    }
    ex.maxM = std::max(ex.maxM, M);
    ex.maxN = std::max(ex.maxN, N);
    for (size_t sgi = 0; sgi < np; ++sgi) {
      const Block& ab = A.blocks[prs[sgi].first];
      const Block& bb = B.blocks[prs[sgi].second];
      GemmSeg sg;
      std::memset(&sg, 0, sizeof(sg));
      sg.a_off = ab.offset;
      sg.b_off = bb.offset;
      for (int l = 0; l < it.mrank; ++l) sg.a_mstr[l] = (uint32_t)am[sgi][l];
      for (int l = 0; l < it.nrank; ++l) sg.b_nstr[l] = (uint32_t)bn[sgi][l];
      // contracted legs
      SmallVec<int64_t> kd, ka, kb;
      for (int i = 0; i < nCA; ++i) {
        kd.push_back(ab.dims[st.cindA[i]]);
        ka.push_back(ab.strides[st.cindA[i]]);
        kb.push_back(bb.strides[st.cindB[i]]);
      }
      merge_legs(kd, ka, &kb, 1);
      int64_t K = 1;
      for (auto d : kd) K *= d;
      if (K >= (1ll << 31)) {
        err = "contracted extent exceeds 2^31";
        return 1;
      }
      sg.K = (int32_t)K;
      sg.koff_idx = (int32_t)ex.koff.size();
      ex.koff.resize(ex.koff.size() + 2 * K);
      for (int64_t q = 0; q < K; ++q) {
        int64_t idx = q, oa = 0, ob = 0;
        for (size_t l = 0; l < kd.size(); ++l) {
          const int64_t i = idx % kd[l];
          idx /= kd[l];
          oa += i * ka[l];
          ob += i * kb[l];
        }
        ex.koff[sg.koff_idx + q] = (uint32_t)oa;
        ex.koff[sg.koff_idx + K + q] = (uint32_t)ob;
      }
      // unit-stride leg decides the load direction
      int64_t minm = INT64_MAX, mink = INT64_MAX, minn = INT64_MAX, minkb = INT64_MAX;
      for (int l = 0; l < it.mrank; ++l) minm = std::min(minm, am[sgi][l]);
      for (int l = 0; l < it.nrank; ++l) minn = std::min(minn, bn[sgi][l]);
      for (size_t l = 0; l < kd.size(); ++l) {
        mink = std::min(mink, ka[l]);
        minkb = std::min(minkb, kb[l]);
      }
      sg.a_kfast = (mink < minm) ? 1 : 0;
      sg.b_kfast = (minkb < minn) ? 1 : 0;
      ex.segs.push_back(sg);
      ex.maxK = std::max(ex.maxK, K);
      ex.macs += (double)M * (double)K * (double)N;
    }
    it.seg_count = (int)ex.segs.size() - it.seg_begin;
    ex.items.push_back(it);
  }
  return 0;
}

// Output blocks grouped by the large operand's open sector tuple (the
// blocks of one group read the same rows of the large operand), and the lean
// records of the per-block streaming kernel where they apply.
static int build_stream_groups(StepExec& ex, const TStruct& C, int bigside, int nOA, int nOB,
                               const std::vector<int>& pos_of_out, int64_t small_total,
                               std::string& err) {
  std::unordered_map<SecKey, int, SecHash> pid;
  std::vector<std::vector<int32_t>> groups;
  SecKey key;
  for (int cb = 0; cb < (int)C.blocks.size(); ++cb) {
    const Block& b = C.blocks[cb];
    if (bigside == 1) {
      key.assign(nOA, 0);
      for (int p = 0; p < nOA; ++p) key[p] = b.sec[pos_of_out[p]];
    } else {
      key.assign(nOB, 0);
      for (int p = 0; p < nOB; ++p) key[p] = b.sec[pos_of_out[nOA + p]];
    }
    auto f = pid.find(key);
    int id;
    if (f == pid.end()) {
      id = (int)groups.size();
      pid.emplace(key, id);
      groups.emplace_back();
    } else {
      id = f->second;
    }
    groups[id].push_back(cb);
  }
  for (const auto& grp : groups) {
    const int J0 = bigside == 1 ? ex.items[grp[0]].M : ex.items[grp[0]].N;
    for (int cb : grp) {
      if ((bigside == 1 ? ex.items[cb].M : ex.items[cb].N) != J0) {
        err = "output blocks of one open sector tuple differ in extent";
        return 1;
      }
    }
    ex.spanel.push_back((int32_t)ex.sorder.size());
    ex.sorder.insert(ex.sorder.end(), grp.begin(), grp.end());
  }
  ex.spanel.push_back((int32_t)ex.sorder.size());
  ex.bstream = bigside;
  ex.small_total = small_total;
  build_lean(ex);
  return 0;
}

// MFMA orientation decided by the largest output block's unit-stride leg.
static void auto_orientation(StepExec& ex) {
  if (ex.items.empty()) return;
  size_t big = 0;
  for (size_t i = 1; i < ex.items.size(); ++i)
    if ((int64_t)ex.items[i].M * ex.items[i].N > (int64_t)ex.items[big].M * ex.items[big].N) big = i;
  const GemmItem& it = ex.items[big];
  uint32_t mm = UINT32_MAX, mn = UINT32_MAX;
  for (int l = 0; l < it.mrank; ++l) mm = std::min(mm, it.c_mstr[l]);
  for (int l = 0; l < it.nrank; ++l) mn = std::min(mn, it.c_nstr[l]);
  ex.tr_auto = mm < mn;
  ex.tr = ex.tr_auto;
}

static int compile_step(const graft_pair_step& st, const TStruct& A, const TStruct& B,
                        const TStruct* fixed_out, TStruct& C, StepExec& ex, std::string& err,
                        int fuse_mode, bool want_holes_a, bool want_holes_b) {
  const int nOA = st.n_oindA, nCA = st.n_cindA, nOB = st.n_oindB, nCB = st.n_cindB;
  if (nCA != nCB || nOA + nCA != A.rank || nOB + nCB != B.rank) {
    err = "pair step leg counts do not match operand ranks";
    return 1;
  }
  const int nOut = st.n_out_codomain + st.n_out_domain;
  if (nOut != nOA + nOB) {
    err = "output partition does not cover the open legs";
    return 1;
  }
  std::vector<int> seenA(A.rank, 0), seenB(B.rank, 0), seenO(nOut, 0);
  for (int i = 0; i < nOA; ++i) seenA.at(st.oindA[i])++;
  for (int i = 0; i < nCA; ++i) seenA.at(st.cindA[i])++;
  for (int i = 0; i < nOB; ++i) seenB.at(st.oindB[i])++;
  for (int i = 0; i < nCB; ++i) seenB.at(st.cindB[i])++;
  for (int i = 0; i < nOut; ++i) seenO.at(st.out[i])++;
  for (int v : seenA) if (v != 1) { err = "A legs are not a partition"; return 1; }
  for (int v : seenB) if (v != 1) { err = "B legs are not a partition"; return 1; }
  for (int v : seenO) if (v != 1) { err = "output permutation is not a permutation"; return 1; }
  for (int i = 0; i < nCA; ++i) {
    if (A.secdims[st.cindA[i]] != B.secdims[st.cindB[i]]) {
      err = "contracted legs have different dimensions or sectors";
      return 1;
    }
  }
  // Output leg q takes intermediate position out[q]; position p < nOA is A leg oindA[p].
  C.rank = nOut;
  C.secdims.assign(nOut, {});
  std::vector<int> pos_of_out(nOut);  // inverse permutation: intermediate pos -> output leg
  for (int q = 0; q < nOut; ++q) {
    const int p = st.out[q];
    pos_of_out[p] = q;
    C.secdims[q] = p < nOA ? A.secdims[st.oindA[p]] : B.secdims[st.oindB[p - nOA]];
  }
  static const bool timing = std::getenv("GRAFT_PLAN_TIMING") != nullptr;
  auto clk = [] { return std::chrono::steady_clock::now(); };
  auto ms_since = [](std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
  };
  auto t0 = clk();
  // Enumerate contributions with integer ids of the sector tuples: r (A's
  // open legs), c (contracted legs), s (B's open legs).  A pair (A block,
  // B block) with the same c feeds output block (r, s).  The pairs of each
  // output block keep enumeration order (A blocks ascending, then B blocks
  // ascending): that is the block's summation order.
  struct TupleIds {
    std::unordered_map<SecKey, int32_t, SecHash> id;
    std::vector<SecKey> keys;
    int32_t get(const SecKey& k) {
      auto f = id.find(k);
      if (f != id.end()) return f->second;
      const int32_t v = (int32_t)keys.size();
      id.emplace(k, v);
      keys.push_back(k);
      return v;
    }
    int32_t find(const SecKey& k) const {
      auto f = id.find(k);
      return f == id.end() ? -1 : f->second;
    }
  };
  TupleIds Rt, Ct, St;
  const int nAb = (int)A.blocks.size(), nBb = (int)B.blocks.size();
  std::vector<int32_t> aR(nAb, -1), aC(nAb), bS(nBb, -1);
  SecKey key;
  for (int ai = 0; ai < nAb; ++ai) {
    key.assign(nCA, 0);
    for (int i = 0; i < nCA; ++i) key[i] = A.blocks[ai].sec[st.cindA[i]];
    aC[ai] = Ct.get(key);
  }
  std::vector<std::vector<int32_t>> bByC(Ct.keys.size());
  for (int bi = 0; bi < nBb; ++bi) {
    key.assign(nCB, 0);
    for (int i = 0; i < nCB; ++i) key[i] = B.blocks[bi].sec[st.cindB[i]];
    const int32_t c = Ct.find(key);
    if (c < 0) continue;
    key.assign(nOB, 0);
    for (int p = 0; p < nOB; ++p) key[p] = B.blocks[bi].sec[st.oindB[p]];
    bS[bi] = St.get(key);
    bByC[c].push_back(bi);
  }
  for (int ai = 0; ai < nAb; ++ai) {
    if (bByC[aC[ai]].empty()) continue;
    key.assign(nOA, 0);
    for (int p = 0; p < nOA; ++p) key[p] = A.blocks[ai].sec[st.oindA[p]];
    aR[ai] = Rt.get(key);
  }
  const uint64_t ns64 = std::max<uint64_t>(1, St.keys.size());
  std::unordered_map<uint64_t, int32_t> kid;   // r * ns + s -> key index
  std::vector<uint64_t> kval;                  // key index -> r * ns + s
  std::vector<int32_t> pa, pb, pk;             // pairs in enumeration order
  kid.reserve(2 * (size_t)nAb + 16);
  for (int ai = 0; ai < nAb; ++ai) {
    if (aR[ai] < 0) continue;
    for (int bi : bByC[aC[ai]]) {
      const uint64_t k = (uint64_t)aR[ai] * ns64 + (uint64_t)bS[bi];
      auto f = kid.find(k);
      int32_t id;
      if (f == kid.end()) {
        id = (int32_t)kval.size();
        kid.emplace(k, id);
        kval.push_back(k);
      } else {
        id = f->second;
      }
      pa.push_back(ai);
      pb.push_back(bi);
      pk.push_back(id);
    }
  }
  const size_t nkeys = kval.size();
  std::vector<SecKey> ksec(nkeys);   // output sector tuple of every key
  for (size_t i = 0; i < nkeys; ++i) {
    const SecKey& rk = Rt.keys[kval[i] / ns64];
    const SecKey& sk = St.keys[kval[i] % ns64];
    SecKey& cs = ksec[i];
    cs.assign(nOut, 0);
    for (int p = 0; p < nOA; ++p) cs[pos_of_out[p]] = rk[p];
    for (int p = 0; p < nOB; ++p) cs[pos_of_out[nOA + p]] = sk[p];
  }
  const double t_contrib = timing ? ms_since(t0) : 0.0;
  auto t1 = clk();
  // Output blocks.
  std::vector<int32_t> kblk(nkeys);   // key -> output block
  if (fixed_out != nullptr) {
    if (fixed_out->rank != nOut) {
      err = "output descriptor rank does not match the plan";
      return 1;
    }
    for (int q = 0; q < nOut; ++q)
      if (fixed_out->secdims[q] != C.secdims[q]) {
        err = "output descriptor legs do not match the plan";
        return 1;
      }
    C = *fixed_out;
    for (size_t i = 0; i < nkeys; ++i) {
      auto f = C.index.find(ksec[i]);
      if (f == C.index.end()) {
        err = "contraction produces a block outside the output structure";
        return 1;
      }
      kblk[i] = f->second;
    }
  } else {
    std::vector<int32_t> order(nkeys);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(),
              [&](int32_t x, int32_t y) { return ksec[x] < ksec[y]; });
    int64_t off = 0;
    C.blocks.clear();
    C.index.clear();
    C.blocks.reserve(nkeys);
    C.index.reserve(nkeys);
    for (size_t j = 0; j < nkeys; ++j) {
      const SecKey& cs = ksec[order[j]];
      Block blk;
      blk.sec = cs;
      blk.dims.resize(nOut);
      blk.strides.resize(nOut);
      int64_t sz = 1;
      for (int q = 0; q < nOut; ++q) {
        blk.dims[q] = C.secdims[q][cs[q]];
        blk.strides[q] = sz;
        sz *= blk.dims[q];
      }
      blk.size = sz;
      blk.offset = off;
      off += sz;
      kblk[order[j]] = (int32_t)j;
      C.index.emplace(blk.sec, (int)C.blocks.size());
      C.blocks.push_back(std::move(blk));
    }
    C.total = off;
    if (C.blocks.empty()) {
      err = "contraction step has no matching blocks";
      return 1;
    }
  }
  // pairs grouped per output block, enumeration order kept inside a block
  std::vector<int64_t> pstart(C.blocks.size() + 1, 0);
  for (int32_t k : pk) pstart[kblk[k] + 1]++;
  for (size_t b = 0; b < C.blocks.size(); ++b) pstart[b + 1] += pstart[b];
  std::vector<std::pair<int, int>> pairs(pk.size());
  {
    std::vector<int64_t> fill(pstart.begin(), pstart.end() - 1);
    for (size_t i = 0; i < pk.size(); ++i) pairs[fill[kblk[pk[i]]]++] = {pa[i], pb[i]};
  }
  const double t_cblocks = timing ? ms_since(t1) : 0.0;
  if ((uint64_t)A.total >= (1ull << 32) - 2 || (uint64_t)B.total >= (1ull << 32) - 2 ||
      (uint64_t)C.total >= (1ull << 32) - 2) {
    err = "tensor exceeds 2^32 elements (32-bit tile offsets)";
    return 1;
  }
  ex.a = st.a;
  ex.b = st.b;
  ex.dst = st.dst;
  ex.conjA = st.conjA;
  ex.conjB = st.conjB;
  ex.items.clear();
  ex.segs.clear();
  ex.koff.clear();
  ex.macs = 0.0;
  // Per-block streaming: in every output block the smaller operand
  // contributes at most kBlockStreamS open elements (TTNO absorbs), so the
  // step is HBM-bound and runs best where the blocks lie.
  int bigside = 0;
  int64_t maxS = 0;
  if (C.blocks.size() >= 2) {
    bigside = A.total >= B.total ? 1 : 2;
    for (const auto& b : C.blocks) {
      int64_t sz = 1;
      if (bigside == 1)
        for (int p = 0; p < nOB; ++p) sz *= b.dims[pos_of_out[nOA + p]];
      else
        for (int p = 0; p < nOA; ++p) sz *= b.dims[pos_of_out[p]];
      maxS = std::max(maxS, sz);
    }
  }
  // Automatic mode streams only when the output blocks sum few block pairs
  // (at most 4 on average): a block that sums dozens re-reads its rows once
  // per pair, and the per-pair records make large plans slow to build; such
  // steps run as fused per-charge GEMMs instead (measured on Hubbard sweeps:
  // D=128, two sweeps 2.9 s against 3.8 s).
  const bool stream_first =
      bigside != 0 &&
      (fuse_mode == 3 || (fuse_mode == 1 && maxS <= kBlockStreamS &&
                          pairs.size() <= 4 * C.blocks.size()));
  // Fused sector path for steps made of many small blocks (decided from the
  // output blocks alone, before any per-block GEMM item is built).
  ex.fused = false;
  ex.bstream = 0;
  ex.bstream_pref = false;
  ex.sorder.clear();
  ex.spanel.clear();
  if (!stream_first && (fuse_mode == 1 || fuse_mode == 2) && C.blocks.size() >= 2 &&
      canonical_blocks(A) && canonical_blocks(B) && canonical_blocks(C)) {
    const double avg = (double)C.total / (double)C.blocks.size();
    // fused when the output blocks are small, or when each sums several block
    // pairs (per-block items would repeat the output tile once per pair and
    // carry one record per pair: measured slower in sweeps, and plans with
    // hundreds of thousands of pairs take tens of ms to build and upload)
    const bool many_pairs = pairs.size() >= 4 * C.blocks.size();
    if (fuse_mode == 2 || ((avg < 64.0 * 64.0 || many_pairs) && C.blocks.size() >= 8)) {
      double macs = 0.0;
      for (size_t cb = 0; cb < C.blocks.size(); ++cb) {
        for (int64_t i = pstart[cb]; i < pstart[cb + 1]; ++i) {
          const Block& ab = A.blocks[pairs[i].first];
          double K = 1.0;
          for (int q = 0; q < nCA; ++q) K *= (double)ab.dims[st.cindA[q]];
          macs += (double)C.blocks[cb].size * K;
        }
      }
      auto t2 = clk();
      if (build_fused(st, A, B, C, pos_of_out, ex, err, want_holes_a, want_holes_b)) return 1;
      if (timing)
        std::fprintf(stderr, "  contrib %.2f ms (%zu keys), C blocks %.2f ms, fused build %.2f ms\n",
                     t_contrib, nkeys, t_cblocks, ms_since(t2));
      ex.macs = macs;   // algorithmic (sector) work, not the padded fused work
      return 0;
    }
  }
  auto t_items = clk();
  if (build_items(st, A, B, C, pos_of_out, pairs, pstart, ex, err)) return 1;
  const double ms_items = timing ? ms_since(t_items) : 0.0;
  auto t_group = clk();
  // Output blocks grouped by the large operand's open sector tuple (the
  // blocks of one group read the same rows of the large operand).
  if (bigside != 0) {
    if (build_stream_groups(ex, C, bigside, nOA, nOB, pos_of_out,
                            bigside == 1 ? B.total : A.total, err))
      return 1;
    ex.bstream_pref = stream_first;
  }
  if (timing)
    std::fprintf(stderr, "  contrib %.2f ms (%zu keys), C blocks %.2f ms, items %.2f ms (%zu segs, koff %zu), "
                 "groups+lean %.2f ms\n", t_contrib, nkeys, t_cblocks, ms_items, ex.segs.size(),
                 ex.koff.size(), ms_since(t_group));
  auto_orientation(ex);
  // Streaming-kernel eligibility: one dense block, short K, one small open side.
  ex.skinny = 0;
  if (ex.items.size() == 1 && ex.segs.size() == 1) {
    const GemmItem& it = ex.items[0];
    const int64_t K = ex.segs[0].K;
    auto fits = [&](int64_t S) {
      return (K <= kSkinnyMaxK && S <= kSkinnyMaxS) ||
             (K <= kSkinnyAccMaxK && S <= kSkinnyAccMaxS);
    };
    if (fits(it.N) && it.M > it.N)
      ex.skinny = 1;
    else if (fits(it.M) && it.N > it.M)
      ex.skinny = 2;
  }
  return 0;
}

// A coupled step: the caller lists every contributing block pair of every
// output block with a real coefficient (non-abelian tensors: the recoupling
// coefficient of the pair's fusion trees).  Blocks are addressed by their
// position in the descriptors; a pair's A and B blocks must agree on the
// contracted legs' sectors, and its output block must carry the pair's open
// sectors.  The pairs of one output block are summed in the order given.
static int compile_coupled_step(const graft_pair_step& st, const TStruct& A, const TStruct& B,
                                const TStruct& C, const int32_t* triples, const double* coefs,
                                int64_t npairs, StepExec& ex, std::string& err) {
  const int nOA = st.n_oindA, nCA = st.n_cindA, nOB = st.n_oindB, nCB = st.n_cindB;
  if (nCA != nCB || nOA + nCA != A.rank || nOB + nCB != B.rank) {
    err = "pair step leg counts do not match operand ranks";
    return 1;
  }
  const int nOut = st.n_out_codomain + st.n_out_domain;
  if (nOut != nOA + nOB || C.rank != nOut) {
    err = "output legs do not match the open legs";
    return 1;
  }
  std::vector<int> seenA(A.rank, 0), seenB(B.rank, 0), seenO(nOut, 0);
  for (int i = 0; i < nOA; ++i) seenA.at(st.oindA[i])++;
  for (int i = 0; i < nCA; ++i) seenA.at(st.cindA[i])++;
  for (int i = 0; i < nOB; ++i) seenB.at(st.oindB[i])++;
  for (int i = 0; i < nCB; ++i) seenB.at(st.cindB[i])++;
  for (int i = 0; i < nOut; ++i) seenO.at(st.out[i])++;
  for (int v : seenA) if (v != 1) { err = "A legs are not a partition"; return 1; }
  for (int v : seenB) if (v != 1) { err = "B legs are not a partition"; return 1; }
  for (int v : seenO) if (v != 1) { err = "output permutation is not a permutation"; return 1; }
  for (int i = 0; i < nCA; ++i)
    if (A.secdims[st.cindA[i]] != B.secdims[st.cindB[i]]) {
      err = "contracted legs have different degeneracies";
      return 1;
    }
  std::vector<int> pos_of_out(nOut);
  for (int q = 0; q < nOut; ++q) {
    const int p = st.out[q];
    pos_of_out[p] = q;
    const auto& want = p < nOA ? A.secdims[st.oindA[p]] : B.secdims[st.oindB[p - nOA]];
    if (C.secdims[q] != want) {
      err = "output leg " + std::to_string(q) + " does not match its open leg";
      return 1;
    }
  }
  if (C.blocks.empty()) {
    err = "coupled step without output blocks";
    return 1;
  }
  const int64_t nA = (int64_t)A.blocks.size(), nB = (int64_t)B.blocks.size();
  const int64_t nC = (int64_t)C.blocks.size();
  std::vector<int64_t> pstart(nC + 1, 0);
  for (int64_t i = 0; i < npairs; ++i) {
    const int32_t c = triples[3 * i], a = triples[3 * i + 1], b = triples[3 * i + 2];
    if (c < 0 || c >= nC || a < 0 || a >= nA || b < 0 || b >= nB) {
      err = "block pair " + std::to_string(i) + " indexes a block out of range";
      return 1;
    }
    const Block& ab = A.blocks[a];
    const Block& bb = B.blocks[b];
    const Block& cb = C.blocks[c];
    for (int q = 0; q < nCA; ++q)
      if (ab.sec[st.cindA[q]] != bb.sec[st.cindB[q]]) {
        err = "block pair " + std::to_string(i) + " disagrees on a contracted sector";
        return 1;
      }
    for (int p = 0; p < nOA; ++p)
      if (cb.sec[pos_of_out[p]] != ab.sec[st.oindA[p]]) {
        err = "block pair " + std::to_string(i) + " does not carry its output block's sectors";
        return 1;
      }
    for (int p = 0; p < nOB; ++p)
      if (cb.sec[pos_of_out[nOA + p]] != bb.sec[st.oindB[p]]) {
        err = "block pair " + std::to_string(i) + " does not carry its output block's sectors";
        return 1;
      }
    if (!std::isfinite(coefs[i])) {
      err = "block pair " + std::to_string(i) + " has a non-finite coefficient";
      return 1;
    }
    pstart[c + 1]++;
  }
  for (int64_t c = 0; c < nC; ++c) pstart[c + 1] += pstart[c];
  std::vector<std::pair<int, int>> pairs((size_t)npairs);
  std::vector<double> cf((size_t)npairs);
  {
    std::vector<int64_t> fill(pstart.begin(), pstart.end() - 1);
    for (int64_t i = 0; i < npairs; ++i) {
      const int64_t j = fill[triples[3 * i]]++;
      pairs[j] = {triples[3 * i + 1], triples[3 * i + 2]};
      cf[j] = coefs[i];
    }
  }
  if ((uint64_t)A.total >= (1ull << 32) - 2 || (uint64_t)B.total >= (1ull << 32) - 2 ||
      (uint64_t)C.total >= (1ull << 32) - 2) {
    err = "tensor exceeds 2^32 elements (32-bit tile offsets)";
    return 1;
  }
  ex.a = st.a;
  ex.b = st.b;
  ex.dst = st.dst;
  ex.conjA = st.conjA;
  ex.conjB = st.conjB;
  ex.items.clear();
  ex.segs.clear();
  ex.koff.clear();
  ex.macs = 0.0;
  ex.fused = false;
  ex.bstream = 0;
  ex.bstream_pref = false;
  ex.skinny = 0;
  ex.lean = false;
  if (build_items(st, A, B, C, pos_of_out, pairs, pstart, ex, err)) return 1;
  ex.coef = std::move(cf);
  // Per-block streaming where one operand adds few open elements to every
  // output block (non-abelian steps are mostly such: many small block pairs
  // with short contractions, latency-bound on the tile kernels).
  if (C.blocks.size() >= 2) {
    const int bigside = A.total >= B.total ? 1 : 2;
    int64_t maxS = 0;
    for (const auto& it : ex.items) maxS = std::max<int64_t>(maxS, bigside == 1 ? it.N : it.M);
    if (build_stream_groups(ex, C, bigside, nOA, nOB, pos_of_out,
                            bigside == 1 ? B.total : A.total, err))
      return 1;
    ex.bstream_pref = maxS <= kBlockStreamS;
  }
  auto_orientation(ex);
  return 0;
}

// Target number of workgroups for one launch (two per CU on 256 CUs).
static constexpr int kTargetBlocks = 512;
// Coupled steps: output tiles wanted before a block's pair list is cut, and
// the most chunks per block.
static constexpr int kCoupledTargetTiles = 4096;
static constexpr int kCoupledMaxSlices = 16;

static int default_config(const StepExec& ex) {
  if (ex.bstream_pref) return kBlockStreamConfig;
  return choose_config(ex.maxM, ex.maxN, ex.skinny);
}

static void assign_tiles(StepExec& ex) {
  // coupled steps: start from the per-block items (the streaming kernels and
  // their work lists index them by output block)
  ex.cslices = 1;
  if (!ex.coef.empty()) {
    if (ex.items_base.empty()) ex.items_base = ex.items;
    else ex.items = ex.items_base;
    ex.xcd_config = -1;
  }
  if (ex.config == kBlockStreamConfig) {
    // One output block per entry while that keeps the wave count moderate
    // (small steps are latency-bound: they need every wave they can get);
    // beyond it an entry walks up to kBlockStreamPanel blocks of one group.
    // Rows outer, blocks inner: neighbouring waves read the same rows.
    int64_t single = 0;
    for (size_t g = 0; g + 1 < ex.spanel.size(); ++g) {
      const GemmItem& it0 = ex.items[ex.sorder[ex.spanel[g]]];
      const int64_t J = ex.bstream == 1 ? it0.M : it0.N;
      single += ((J + kBlockStreamRows - 1) / kBlockStreamRows) * (ex.spanel[g + 1] - ex.spanel[g]);
    }
    const int panel = (int)std::min<int64_t>(
        kBlockStreamPanel, std::max<int64_t>(1, (single + kBlockStreamWaves - 1) / kBlockStreamWaves));
    ex.swork.clear();
    for (size_t g = 0; g + 1 < ex.spanel.size(); ++g) {
      const int b = ex.spanel[g], e = ex.spanel[g + 1];
      const GemmItem& it0 = ex.items[ex.sorder[b]];
      const int J = ex.bstream == 1 ? it0.M : it0.N;
      for (int j0 = 0; j0 < J; j0 += kBlockStreamRows)
        for (int c0 = b; c0 < e; c0 += panel)
          ex.swork.push_back({c0, std::min(panel, e - c0), j0, 0});
    }
    ex.total_tiles = (int)ex.swork.size();
    ex.ksplit = 1;
    ex.partial_bytes = 0;
    return;
  }
  if (ex.config == kSkinnyConfig) {
    const GemmItem& it = ex.items[0];
    ex.total_tiles = skinny_grid(ex.skinny == 1 ? it.M : it.N);
    ex.ksplit = 1;
    ex.partial_bytes = 0;
    return;
  }
  const TileConfig& c = kTileConfigs[ex.config];
  int t = 0;
  for (auto& it : ex.items) {
    it.tile_begin = t;
    it.tiles_n = (it.N + c.bn - 1) / c.bn;
    t += ((it.M + c.bm - 1) / c.bm) * it.tiles_n;
  }
  ex.total_tiles = t;
  if (!ex.coef.empty() && t > 0 && t < kCoupledTargetTiles) {
    // A coupled output block sums its block pairs one after another in one
    // workgroup: with few output tiles the longest pair list sets the time.
    // Cut every block's list into S chunks of equal k work (k-tiles + 1 per
    // pair); chunk s of every block forms the s-th copy of the item list,
    // storing into slice s of the partial buffer.
    int maxseg = 0;
    int64_t ctot = 0;
    for (const auto& it : ex.items) {
      maxseg = std::max(maxseg, it.seg_count);
      ctot = std::max<int64_t>(ctot, it.c_off + (int64_t)it.M * it.N);
    }
    const int S = (int)std::min<int64_t>({(kCoupledTargetTiles + t - 1) / t, (int64_t)maxseg,
                                          (int64_t)kCoupledMaxSlices});
    if (S >= 2) {
      std::vector<GemmItem> cut;
      cut.reserve(ex.items.size() * S);
      std::vector<int32_t> bounds(S + 1);
      for (int sl = 0; sl < S; ++sl) {
        for (const auto& it : ex.items) {
          // chunk boundaries of this block: first pair whose preceding work
          // reaches sl / S of the total (recomputed per slice: blocks are few)
          int64_t W = 0;
          for (int q = it.seg_begin; q < it.seg_begin + it.seg_count; ++q)
            W += (ex.segs[q].K + c.bk - 1) / c.bk + 1;
          int32_t lo = it.seg_begin + it.seg_count, hi = lo;
          int64_t cum = 0;
          for (int q = it.seg_begin; q < it.seg_begin + it.seg_count; ++q) {
            const int64_t w = (ex.segs[q].K + c.bk - 1) / c.bk + 1;
            if (lo == it.seg_begin + it.seg_count && cum * S >= W * sl) lo = q;
            if (cum * S >= W * (sl + 1)) { hi = q; break; }
            cum += w;
          }
          if (sl == S - 1) hi = it.seg_begin + it.seg_count;
          GemmItem ci = it;
          ci.c_off = it.c_off + (int64_t)sl * ctot;
          ci.seg_begin = lo;
          ci.seg_count = std::max(0, hi - lo);
          cut.push_back(ci);
        }
      }
      ex.items = std::move(cut);
      t = 0;
      for (auto& it : ex.items) {
        it.tile_begin = t;
        t += ((it.M + c.bm - 1) / c.bm) * it.tiles_n;
      }
      ex.total_tiles = t;
      ex.cslices = S;
      ex.ctot = ctot;
    }
  }
  // Split K when one output block has too few tiles to fill the chip and a
  // long contraction (environment closures contract D*D*d legs at once).
  ex.ksplit = 1;
  ex.partial_bytes = ex.cslices > 1 ? (uint64_t)ex.cslices * ex.ctot * 16 : 0;
  if (ex.items.size() == 1 && ex.segs.size() == 1 && t > 0 && t < kTargetBlocks &&
      ex.coef.empty()) {   // coupled steps run whole tiles (mapped launch)
    const int64_t K = ex.segs[0].K;
    const int64_t ktiles = (K + c.bk - 1) / c.bk;
    int64_t ks = (kTargetBlocks + t - 1) / t;
    ks = std::min<int64_t>(ks, ktiles / 8);  // keep at least 8 k-tiles per chunk
    if (ks >= 2) {
      ex.ksplit = (int)ks;
      ex.partial_bytes = (uint64_t)ks * ex.items[0].M * ex.items[0].N * 16;
    }
  }
}

static void update_workspace_size(graft_plan_impl& p) {
  uint64_t part = 0, fused = 0, planes = 0;
  for (auto& ex : p.steps) {
    part = std::max(part, ex.partial_bytes);
    fused = std::max(fused, ex.fused_bytes);
    planes = std::max(planes, ex.bplane_bytes);
  }
  p.ws_partial = ((p.ws_slots + 255) / 256) * 256;
  p.ws_fused = p.ws_partial + ((part + 255) / 256) * 256;
  p.ws_planes = p.ws_fused + ((fused + 255) / 256) * 256;
  p.ws_bytes = p.ws_planes + planes;
}

// Pinned staging buffer for plan metadata uploads (grown on demand, kept
// until the process ends): a pageable copy of tens of MB costs several times
// the pinned one, and sweeps compile a plan for nearly every bond update.
static std::mutex g_stage_mu;
static char* g_stage = nullptr;
static size_t g_stage_bytes = 0;

static int upload_step_meta(graft_plan_impl& p) {
  // one allocation: per step [items | segs | koff], 256-byte aligned pieces
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t total = 0;
  for (auto& ex : p.steps) {
    total += al(ex.items.size() * sizeof(GemmItem));
    total += al(std::max<size_t>(1, ex.segs.size()) * sizeof(GemmSeg));
    total += al(std::max<size_t>(1, ex.koff.size()) * sizeof(uint32_t));
    for (auto* v : {&ex.pa, &ex.pb, &ex.pc}) total += al(v->size() * sizeof(PermBlock));
    for (auto* v : {&ex.ca, &ex.cb, &ex.cc}) total += al(v->size() * sizeof(int32_t));
    total += al(ex.sorder.size() * sizeof(int32_t)) + al(ex.swork.size() * sizeof(StreamWork));
    total += al(ex.litems.size() * sizeof(LeanItem)) + al(ex.lsegs.size() * sizeof(LeanSeg)) +
             al(ex.lwork.size() * sizeof(LeanWork));
    total += al(ex.holes_a.size() * sizeof(HoleRect)) + al(ex.holes_b.size() * sizeof(HoleRect));
    total += al(ex.coef.size() * sizeof(double));
  }
  if (p.d_meta) {
    (void)hipFree(p.d_meta);
    p.d_meta = nullptr;
  }
  GRAFT_CHECK_HIP(hipMalloc(&p.d_meta, total));
  p.meta_bytes = total;
  std::lock_guard<std::mutex> stage_lock(g_stage_mu);
  if (g_stage_bytes < total) {
    if (g_stage) (void)hipHostFree(g_stage);
    g_stage = nullptr;
    g_stage_bytes = 0;
    const size_t want = std::max<size_t>(total, 16u << 20);
    GRAFT_CHECK_HIP(hipHostMalloc((void**)&g_stage, want, hipHostMallocDefault));
    g_stage_bytes = want;
  }
  char* host = g_stage;   // padding bytes between pieces are never read
  size_t off = 0;
  for (auto& ex : p.steps) {
    size_t s1 = ex.items.size() * sizeof(GemmItem);
    std::memcpy(host + off, ex.items.data(), s1);
    ex.d_items = (GemmItem*)((char*)p.d_meta + off);
    off += al(s1);
    size_t s2 = ex.segs.size() * sizeof(GemmSeg);
    if (s2) std::memcpy(host + off, ex.segs.data(), s2);
    ex.d_segs = (GemmSeg*)((char*)p.d_meta + off);
    off += al(std::max<size_t>(1, ex.segs.size()) * sizeof(GemmSeg));
    size_t s3 = ex.koff.size() * sizeof(uint32_t);
    if (s3) std::memcpy(host + off, ex.koff.data(), s3);
    ex.d_koff = (uint32_t*)((char*)p.d_meta + off);
    off += al(std::max<size_t>(1, ex.koff.size()) * sizeof(uint32_t));
    PermBlock** dp[3] = {&ex.d_pa, &ex.d_pb, &ex.d_pc};
    std::vector<PermBlock>* hp[3] = {&ex.pa, &ex.pb, &ex.pc};
    for (int i = 0; i < 3; ++i) {
      const size_t sz = hp[i]->size() * sizeof(PermBlock);
      if (sz) std::memcpy(host + off, hp[i]->data(), sz);
      *dp[i] = sz ? (PermBlock*)((char*)p.d_meta + off) : nullptr;
      off += al(sz);
    }
    int32_t** dc[3] = {&ex.d_ca, &ex.d_cb, &ex.d_cc};
    std::vector<int32_t>* hc[3] = {&ex.ca, &ex.cb, &ex.cc};
    for (int i = 0; i < 3; ++i) {
      const size_t sz = hc[i]->size() * sizeof(int32_t);
      if (sz) std::memcpy(host + off, hc[i]->data(), sz);
      *dc[i] = sz ? (int32_t*)((char*)p.d_meta + off) : nullptr;
      off += al(sz);
    }
    const size_t so = ex.sorder.size() * sizeof(int32_t);
    if (so) std::memcpy(host + off, ex.sorder.data(), so);
    ex.d_sorder = so ? (int32_t*)((char*)p.d_meta + off) : nullptr;
    off += al(so);
    const size_t sw = ex.swork.size() * sizeof(StreamWork);
    if (sw) std::memcpy(host + off, ex.swork.data(), sw);
    ex.d_swork = sw ? (StreamWork*)((char*)p.d_meta + off) : nullptr;
    off += al(sw);
    auto put = [&](const void* src, size_t sz) -> void* {
      if (!sz) return nullptr;
      std::memcpy(host + off, src, sz);
      void* d = (char*)p.d_meta + off;
      off += al(sz);
      return d;
    };
    ex.d_litems = (LeanItem*)put(ex.litems.data(), ex.litems.size() * sizeof(LeanItem));
    ex.d_lsegs = (LeanSeg*)put(ex.lsegs.data(), ex.lsegs.size() * sizeof(LeanSeg));
    ex.d_lwork = (LeanWork*)put(ex.lwork.data(), ex.lwork.size() * sizeof(LeanWork));
    ex.d_holes_a = (HoleRect*)put(ex.holes_a.data(), ex.holes_a.size() * sizeof(HoleRect));
    ex.d_holes_b = (HoleRect*)put(ex.holes_b.data(), ex.holes_b.size() * sizeof(HoleRect));
    ex.d_coef = (double*)put(ex.coef.data(), ex.coef.size() * sizeof(double));
  }
  GRAFT_CHECK_HIP(hipMemcpy(p.d_meta, host, total, hipMemcpyHostToDevice));
  return GRAFT_OK;
}

// Offset maps of every fused step, built on the device once per plan from
// the block tables (4 bytes per element): for the result, blocked element ->
// fused offset; for the operands the inverse, fused offset -> blocked
// element (kMapZero where no block covers the entry), so that both copy
// directions are gathers with coalesced writes.
static int build_maps(graft_ctx* ctx, graft_plan_impl& p) {
  size_t total = 0;
  int64_t tmp_n = 0;
  auto al = [](int64_t n) { return ((size_t)n * sizeof(uint32_t) + 255) & ~size_t(255); };
  for (auto& ex : p.steps)
    if (ex.fused) {
      total += al(ex.fa) + al(ex.fb) + al(ex.totC) + (ex.c_complete ? al(ex.fc) : 0) +
               (ex.ho_step >= 0 ? al(ex.fc) : 0);
      tmp_n = std::max(tmp_n, std::max(ex.totA, ex.totB));
    }
  if (total == 0) return GRAFT_OK;
  GRAFT_CHECK_HIP(hipMalloc(&p.d_maps, total));
  p.map_bytes = total;
  uint32_t* tmp = nullptr;
  GRAFT_CHECK_HIP(hipMalloc(&tmp, al(tmp_n)));
  char* base = (char*)p.d_maps;
  hipError_t e = hipSuccess;
  for (auto& ex : p.steps) {
    if (!ex.fused || e != hipSuccess) continue;
    ex.d_ma = (uint32_t*)base;
    base += al(ex.fa);
    ex.d_mb = (uint32_t*)base;
    base += al(ex.fb);
    ex.d_mc = (uint32_t*)base;
    base += al(ex.totC);
    ex.d_mci = nullptr;
    if (ex.c_complete) {
      ex.d_mci = (uint32_t*)base;
      base += al(ex.fc);
    }
    ex.d_mho = nullptr;
    if (ex.ho_step >= 0) {
      ex.d_mho = (uint32_t*)base;
      base += al(ex.fc);
    }
    // (an operand without blocks has no forward map: its inverse is all zero)
    const int64_t nA = ex.pa.empty() ? 0 : ex.totA, nB = ex.pb.empty() ? 0 : ex.totB;
    // a producer handing its result to this step stores through its own
    // fused-to-blocked map composed with this step's forward operand map
    auto compose = [&](int from) {
      if (from < 0 || e != hipSuccess) return;
      const StepExec& es = p.steps[from];
      e = launch_compose_map(es.d_mci, es.fc, tmp, es.d_mho, ctx->stream);
    };
    e = launch_perm_map(ex.d_pa, (int)ex.pa.size(), ex.d_ca, nA, tmp, ctx->stream);
    if (e == hipSuccess && ex.ho_in_a) compose(ex.ho_from_a);
    if (e == hipSuccess) e = launch_invert_map(tmp, nA, ex.d_ma, ex.fa, ctx->stream);
    if (e == hipSuccess)
      e = launch_perm_map(ex.d_pb, (int)ex.pb.size(), ex.d_cb, nB, tmp, ctx->stream);
    if (e == hipSuccess && ex.ho_in_b) compose(ex.ho_from_b);
    if (e == hipSuccess) e = launch_invert_map(tmp, nB, ex.d_mb, ex.fb, ctx->stream);
    if (e == hipSuccess)
      e = launch_perm_map(ex.d_pc, (int)ex.pc.size(), ex.d_cc, ex.totC, ex.d_mc, ctx->stream);
    if (e == hipSuccess && ex.d_mci)
      e = launch_invert_map(ex.d_mc, ex.totC, ex.d_mci, ex.fc, ctx->stream);
  }
  // the maps read the block tables, which a later autotune may re-upload
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  GRAFT_CHECK_HIP(e);
  GRAFT_CHECK_HIP(es);
  return GRAFT_OK;
}

static int ensure_workspace(graft_ctx* ctx, uint64_t bytes) {
  if (bytes <= ctx->ws_size) return GRAFT_OK;
  if (ctx->capturing) {
    set_error("the plan needs " + std::to_string(bytes) +
              " workspace bytes but a graph capture is open: execute the plan once before capturing");
    return GRAFT_E_INVALID;
  }
  if (ctx->ws_cap && bytes > ctx->ws_cap) {
    set_error("plan workspace of " + std::to_string(bytes) + " bytes exceeds the cap of " +
              std::to_string(ctx->ws_cap) + " bytes");
    return GRAFT_E_NOMEM;
  }
  GRAFT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->ws) {
    GRAFT_CHECK_HIP(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_size = 0;
  }
  hipError_t e = hipMalloc(&ctx->ws, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ctx->ws = nullptr;
    set_error("workspace allocation of " + std::to_string(bytes) + " bytes failed: " +
              hipGetErrorString(e));
    return GRAFT_E_NOMEM;
  }
  ctx->ws_size = bytes;
  ++ctx->ws_generation;
  return GRAFT_OK;
}

static int run_step(graft_ctx* ctx, graft_plan_impl& p, StepExec& ex, const void* const* inputs,
                    void* output, int precision, char* ws) {
  auto ptr = [&](int slot) -> void* {
    if (slot < p.ninputs) return const_cast<void*>(inputs[slot]);
    if (slot == p.output_slot) return output;
    return ws + p.ws_offset[slot];
  };
  void* A = ptr(ex.a);
  void* B = ptr(ex.b);
  void* C = ptr(ex.dst);
  if (!A || !B || !C) {
    set_error("null operand pointer");
    return GRAFT_E_INVALID;
  }
  void* partial = ex.ksplit > 1 ? ws + p.ws_partial : nullptr;
  void* Cdst = C;
  auto mapped = [&](const StepExec& e) {
    return e.fused && ctx->fused_map_loads && precision == GRAFT_PREC_FP64 && e.d_mci != nullptr &&
           e.config < kNumTileConfigs && e.ksplit == 1;
  };
  if (mapped(ex)) {
    // The GEMM writes the result blocks through the element map (no scatter
    // pass).  An operand the tiles read at most twice is also read through
    // its map; one read more often is gathered once into its fused matrix.
    // Handoffs: an operand the producing step already wrote in this step's
    // fused layout is read in place; a result this step hands to the next
    // fused step is stored in that step's layout (its holes zeroed first).
    const TileConfig& tc = kTileConfigs[ex.config];
    const bool ina = ex.ho_in_a && mapped(p.steps[ex.ho_from_a]);
    const bool inb = ex.ho_in_b && mapped(p.steps[ex.ho_from_b]);
    const bool hout = ex.ho_step >= 0 && mapped(p.steps[ex.ho_step]);
    const bool ma = !ina && (ex.maxN + tc.bn - 1) / tc.bn <= 2;
    const bool mb = !inb && (ex.maxM + tc.bm - 1) / tc.bm <= 2;
    auto al = [](int64_t x) { return (size_t)((x * 16 + 255) / 256) * 256; };
    char* fb = ws + p.ws_fused;
    void* Af = ina ? A : fb;
    void* Bf = inb ? B : fb + al(ex.fa);
    if (!ma && !ina) GRAFT_CHECK_HIP(launch_map_gather(A, Af, ex.d_ma, ex.fa, ctx->stream));
    if (!mb && !inb) GRAFT_CHECK_HIP(launch_map_gather(B, Bf, ex.d_mb, ex.fb, ctx->stream));
    if (hout) {
      const StepExec& et = p.steps[ex.ho_step];
      const HoleRect* h = ex.ho_side == 0 ? et.d_holes_a : et.d_holes_b;
      const int nh = (int)(ex.ho_side == 0 ? et.holes_a.size() : et.holes_b.size());
      GRAFT_CHECK_HIP(launch_zero_holes(h, nh, C, ctx->stream));
    }
    const FusedMaps maps{ma ? ex.d_ma : nullptr, mb ? ex.d_mb : nullptr, hout ? ex.d_mho : ex.d_mci};
    if (ex.config == 11 || ex.config == 15 || ex.config == 14 || ex.config == 18) {
      // every component of a fused step loads A and B in one direction
      const int dirs = ex.segs.empty() ? 0
                                       : (ex.segs[0].a_kfast ? 1 : 0) | (ex.segs[0].b_kfast ? 2 : 0);
      GRAFT_CHECK_HIP(launch_zgemm_dbmap(ex.config, ma ? A : Af, mb ? B : Bf, C, ex.d_items,
                                         (int)ex.items.size(), ex.d_segs, ex.total_tiles,
                                         ex.conjA, ex.conjB, ex.tr, dirs, maps, &xcd_split(ex),
                                         ctx->stream));
      return GRAFT_OK;
    }
    GRAFT_CHECK_HIP(launch_zgemm(ex.config, precision, ma ? A : Af, mb ? B : Bf, C, ex.d_items,
                                 (int)ex.items.size(), ex.d_segs, ex.d_koff, ex.total_tiles,
                                 ex.conjA, ex.conjB, 1, nullptr, ex.tr, ctx->stream, &maps));
    return GRAFT_OK;
  }
  if (ex.d_coef != nullptr && ex.config < kNumTileConfigs) {
    // coupled step: every block pair scaled by its coefficient as A is staged
    if (precision != GRAFT_PREC_FP64) {
      set_error("coupled (non-abelian) steps multiply in FP64 only");
      return GRAFT_E_UNSUPPORTED;
    }
    const FusedMaps maps{nullptr, nullptr, nullptr, ex.d_coef};
    void* dst = ex.cslices > 1 ? (void*)(ws + p.ws_partial) : C;
    GRAFT_CHECK_HIP(launch_zgemm(ex.config, precision, A, B, dst, ex.d_items, (int)ex.items.size(),
                                 ex.d_segs, ex.d_koff, ex.total_tiles, ex.conjA, ex.conjB, 1,
                                 nullptr, ex.tr, ctx->stream, &maps));
    if (ex.cslices > 1)
      GRAFT_CHECK_HIP(launch_sum_slices(dst, ex.cslices, ex.ctot, C, ctx->stream));
    return GRAFT_OK;
  }
  if (ex.fused) {
    // gather both operands into the per-charge matrices, multiply, scatter below
    auto al = [](int64_t x) { return (size_t)((x * 16 + 255) / 256) * 256; };
    char* fb = ws + p.ws_fused;
    void* Af = fb;
    void* Bf = fb + al(ex.fa);
    void* Cf = fb + al(ex.fa) + al(ex.fb);
    GRAFT_CHECK_HIP(launch_map_gather(A, Af, ex.d_ma, ex.fa, ctx->stream));
    GRAFT_CHECK_HIP(launch_map_gather(B, Bf, ex.d_mb, ex.fb, ctx->stream));
    A = Af;
    B = Bf;
    C = Cf;
  }
  // split-word handoff of a BF16X3 apply: decided at run time for both steps
  auto planes_ok = [&](const StepExec& prod, const StepExec& cons) {
    return precision == GRAFT_PREC_BF16X3 && ctx->split_words && prod.config == kSkinnyConfig &&
           !cons.fused &&
           cons.config < kNumTileConfigs && cons.ksplit == 1;
  };
  const bool pl_out = ex.pl_to >= 0 && planes_ok(ex, p.steps[ex.pl_to]);
  if (ex.pl_from >= 0 && planes_ok(p.steps[ex.pl_from], ex)) {
    void* Bw = ws + p.ws_planes;
    GRAFT_CHECK_HIP(launch_split_bplanes(B, (int64_t)(ex.bplane_bytes / 16), Bw, ex.conjB, ctx->stream));
    GRAFT_CHECK_HIP(launch_zgemm_lowp(ex.config, precision, A, Bw, C, ex.d_items, (int)ex.items.size(),
                                      ex.d_segs, ex.d_koff, ex.total_tiles, ex.conjA, 0, 1, nullptr,
                                      ex.tr, ctx->stream, true));
    return GRAFT_OK;
  }
  if (ex.config == kBlockStreamConfig && ex.lean && ctx->lean_stream) {
    // HBM-bound, FP64 whatever the precision (as the streaming kernels below)
    GRAFT_CHECK_HIP(launch_zstream_lean(ex.bstream == 1, ex.lean_smax, A, B, C, ex.d_litems,
                                        ex.d_lsegs, ex.d_lwork, (int)ex.lwork.size(),
                                        (int)std::min<int64_t>(ex.small_total, INT32_MAX), ex.conjA,
                                        ex.conjB, ctx->stream));
  } else if (ex.config == kBlockStreamConfig) {
    // HBM-bound, FP64 whatever the precision (as the streaming kernel below)
    GRAFT_CHECK_HIP(launch_zstream_blocks(ex.bstream == 1, A, B, C, ex.d_items, ex.d_segs,
                                          ex.d_koff, ex.d_sorder, ex.d_swork,
                                          (int)ex.swork.size(), ex.conjA, ex.conjB, ctx->stream,
                                          ex.d_coef));
  } else if (ex.config == kSkinnyConfig) {
    // HBM-bound: the streaming kernel always multiplies in FP64, whatever
    // precision the compute-bound steps use.  One launch per item (a fused
    // step has one item per charge component).
    for (size_t i = 0; i < ex.items.size(); ++i) {
      const GemmItem& it = ex.items[i];
      const GemmSeg& sg = ex.segs[it.seg_begin];
      const int64_t S = ex.skinny == 1 ? it.N : it.M;
      const bool acc = sg.K > kSkinnyMaxK || S > kSkinnyMaxS;
      const int grid = ex.items.size() == 1 ? ex.total_tiles : skinny_grid(ex.skinny == 1 ? it.M : it.N);
      GRAFT_CHECK_HIP(launch_zskinny(ex.skinny == 1, acc, A, B, C, ex.d_items + i,
                                     ex.d_segs + it.seg_begin, ex.d_koff, grid, ex.conjA, ex.conjB,
                                     ctx->stream, pl_out));
    }
  } else {
    const int kfast = precision == GRAFT_PREC_FP64 ? piped_kfast(ex, ex.config) : -1;
    hipError_t e = launch_zgemm(ex.config, precision, A, B, C, ex.d_items, (int)ex.items.size(),
                                ex.d_segs, ex.d_koff, ex.total_tiles, ex.conjA, ex.conjB,
                                ex.ksplit, partial, ex.tr, ctx->stream, nullptr, kfast,
                                &xcd_split(ex));
    if (e == hipErrorNotSupported) {
      set_error("precision not supported by this build");
      return GRAFT_E_UNSUPPORTED;
    }
    GRAFT_CHECK_HIP(e);
  }
  if (ex.fused) GRAFT_CHECK_HIP(launch_map_gather(C, Cdst, ex.d_mc, ex.totC, ctx->stream));
  return GRAFT_OK;
}

static int execute_plan(graft_ctx* ctx, graft_plan_impl& p, const void* const* inputs,
                        void* output, int precision) {
  if (!inputs || !output) {
    set_error("null inputs/output");
    return GRAFT_E_INVALID;
  }
  int rc = set_device(ctx);
  if (rc) return rc;
  rc = ensure_workspace(ctx, p.ws_bytes);
  if (rc) return rc;
  for (auto& ex : p.steps) {
    rc = run_step(ctx, p, ex, inputs, output, precision, (char*)ctx->ws);
    if (rc) return rc;
  }
  return GRAFT_OK;
}

}  // namespace graft

// Workspace of graft_heff_apply_noalloc: a file-scope device array, so that
// entry point never allocates.  Each device holds its own copy from module
// load until the process ends; its address is looked up once per device when
// a context is created there (the array's name is only a host-side handle).
static constexpr uint64_t kNoallocScratchBytes = 128ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
static constexpr int kMaxDevices = 64;
static constexpr int kNoallocMaxInputs = 32;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

static int init_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) {
    graft::set_error("device index out of range for the static scratch table");
    return GRAFT_E_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return GRAFT_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    (void)hipGetLastError();
    graft::set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    return GRAFT_E_HIP;
  }
  g_scratch_ptr[device] = ptr;
  return GRAFT_OK;
}

using namespace graft;

extern "C" {

const char* graft_last_error(void) { return g_last_error.c_str(); }
const char* graft_version(void) { return "graft-heff 0.1 (gfx950)"; }

int graft_ctx_create(int device, void* hip_stream, graft_ctx** out) {
  GRAFT_REQUIRE(out != nullptr, "null output handle");
  std::unique_ptr<graft_ctx> c(new graft_ctx());
  c->device = device;
  c->stream = (hipStream_t)hip_stream;
  GRAFT_CHECK_HIP(hipSetDevice(device));
  const int rs = init_scratch(device);
  if (rs) return rs;
  GRAFT_CHECK_HIP(hipMalloc((void**)&c->d_partials,
                            sizeof(double) * 2 * kReduceBlocks * kMaxMultiVec));
  GRAFT_CHECK_HIP(hipMalloc((void**)&c->d_result, sizeof(double) * 2 * kMaxMultiVec));
  GRAFT_CHECK_HIP(hipHostMalloc((void**)&c->h_result, sizeof(double) * 2 * kMaxMultiVec));
  GRAFT_CHECK_HIP(hipMalloc((void**)&c->d_coef, sizeof(double) * 2 * kMaxMultiVec));
  GRAFT_CHECK_HIP(
      hipMalloc((void**)&c->d_comb, sizeof(double) * 2 * kMaxMultiVec * kMaxCombine));
  *out = c.release();
  return GRAFT_OK;
}

int graft_ctx_destroy(graft_ctx* ctx) {
  if (!ctx) return GRAFT_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->ws) (void)hipFree(ctx->ws);
  (void)hipFree(ctx->d_partials);
  (void)hipFree(ctx->d_result);
  (void)hipHostFree(ctx->h_result);
  (void)hipFree(ctx->d_coef);
  (void)hipFree(ctx->d_comb);
  delete ctx;
  return GRAFT_OK;
}

int graft_ctx_set_stream(graft_ctx* ctx, void* hip_stream) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  GRAFT_REQUIRE(!ctx->capturing || (hipStream_t)hip_stream == ctx->stream,
                "cannot change the stream while a graph capture is open");
  ctx->stream = (hipStream_t)hip_stream;
  return GRAFT_OK;
}

int graft_ctx_set_workspace_cap(graft_ctx* ctx, uint64_t bytes) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  ctx->ws_cap = bytes;
  return GRAFT_OK;
}

int graft_ctx_workspace_bytes(graft_ctx* ctx, uint64_t* bytes) {
  GRAFT_REQUIRE(ctx != nullptr && bytes != nullptr, "null argument");
  *bytes = ctx->ws_size;
  return GRAFT_OK;
}

int graft_ctx_set_sector_fusion(graft_ctx* ctx, int32_t mode) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  GRAFT_REQUIRE(mode >= 0 && mode <= 3, "sector fusion mode must be 0, 1, 2 or 3");
  ctx->sector_fuse = mode;
  return GRAFT_OK;
}

int graft_ctx_set_fused_map_loads(graft_ctx* ctx, int32_t on) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  ctx->fused_map_loads = on ? 1 : 0;
  return GRAFT_OK;
}

int graft_ctx_set_split_words(graft_ctx* ctx, int32_t on) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  ctx->split_words = on ? 1 : 0;
  return GRAFT_OK;
}

int graft_ctx_set_lean_stream(graft_ctx* ctx, int32_t on) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  ctx->lean_stream = on ? 1 : 0;
  return GRAFT_OK;
}

int graft_ctx_reserve_workspace(graft_ctx* ctx, uint64_t bytes) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  int rc = set_device(ctx);
  if (rc) return rc;
  return ensure_workspace(ctx, bytes);
}

int graft_ctx_release_workspace(graft_ctx* ctx) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  int rc = set_device(ctx);
  if (rc) return rc;
  GRAFT_REQUIRE(!ctx->capturing, "cannot release the workspace while a graph capture is open");
  GRAFT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->ws) GRAFT_CHECK_HIP(hipFree(ctx->ws));
  ctx->ws = nullptr;
  ctx->ws_size = 0;
  ++ctx->ws_generation;
  return GRAFT_OK;
}

int graft_synchronize(graft_ctx* ctx) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  GRAFT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return GRAFT_OK;
}

int graft_malloc(graft_ctx* ctx, uint64_t bytes, void** dptr) {
  GRAFT_REQUIRE(ctx != nullptr && dptr != nullptr, "null argument");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(std::string("device allocation failed: ") + hipGetErrorString(e));
    return GRAFT_E_NOMEM;
  }
  return GRAFT_OK;
}

int graft_free(graft_ctx* ctx, void* dptr) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  if (dptr) GRAFT_CHECK_HIP(hipFree(dptr));
  return GRAFT_OK;
}

int graft_upload(graft_ctx* ctx, void* dst, const void* host_src, uint64_t bytes) {
  GRAFT_REQUIRE(ctx != nullptr && dst != nullptr && host_src != nullptr, "null argument");
  GRAFT_CHECK_HIP(hipMemcpyAsync(dst, host_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  GRAFT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return GRAFT_OK;
}

int graft_download(graft_ctx* ctx, void* host_dst, const void* src, uint64_t bytes) {
  GRAFT_REQUIRE(ctx != nullptr && host_dst != nullptr && src != nullptr, "null argument");
  GRAFT_CHECK_HIP(hipMemcpyAsync(host_dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  GRAFT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return GRAFT_OK;
}

// Explicit block pairs of a coupled plan (graft_plan_create_coupled).
struct CoupledSpec {
  const graft_tensor_desc* outs;   // per step: its result's block list
  const int64_t* npairs;           // per step
  const int32_t* pairs;            // (C block, A block, B block) triples, steps concatenated
  const double* coefs;             // one per pair
};

static int plan_create_impl(graft_ctx* ctx, const graft_pair_step* steps, int32_t nsteps,
                            int32_t nslots, int32_t output_slot,
                            const graft_tensor_desc* input_descs, int32_t ninputs,
                            const graft_tensor_desc* out_desc, const CoupledSpec* cp,
                            graft_plan** out) {
  GRAFT_REQUIRE(ctx != nullptr && out != nullptr, "null argument");
  GRAFT_REQUIRE(steps != nullptr && nsteps >= 1, "plan without steps");
  GRAFT_REQUIRE(input_descs != nullptr && ninputs >= 1, "plan without inputs");
  GRAFT_REQUIRE(nslots == ninputs + nsteps, "nslots must equal ninputs + nsteps");
  GRAFT_REQUIRE(output_slot >= ninputs && output_slot < nslots, "output slot out of range");
  int rc = set_device(ctx);
  if (rc) return rc;
  std::unique_ptr<graft_plan> plan(new graft_plan());
  graft_plan_impl& p = plan->impl;
  p.nslots = nslots;
  p.output_slot = output_slot;
  p.ninputs = ninputs;
  p.device = ctx->device;
  p.slots.resize(nslots);
  std::vector<int> defined(nslots, 0), consumed(nslots, 0), born(nslots, -1), dies(nslots, -1);
  std::string err;
  for (int i = 0; i < ninputs; ++i) {
    if (build_struct(input_descs[i], p.slots[i], err, cp != nullptr)) {
      set_error("input " + std::to_string(i) + ": " + err);
      return GRAFT_E_INVALID;
    }
    defined[i] = 1;
  }
  TStruct outfixed;
  bool have_out = false;
  if (out_desc && !cp) {
    if (build_struct(*out_desc, outfixed, err)) {
      set_error("output descriptor: " + err);
      return GRAFT_E_INVALID;
    }
    have_out = true;
  }
  p.steps.resize(nsteps);
  std::vector<int> slot_producer(nslots, -1);
  static const bool timing = std::getenv("GRAFT_PLAN_TIMING") != nullptr;
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto t_start = now();
  int64_t pair_off = 0;
  for (int s = 0; s < nsteps; ++s) {
    const graft_pair_step& st = steps[s];
    auto t_step = now();
    GRAFT_REQUIRE(st.a >= 0 && st.a < nslots && st.b >= 0 && st.b < nslots && st.dst >= 0 &&
                      st.dst < nslots,
                  "step slot out of range");
    GRAFT_REQUIRE(defined[st.a] && defined[st.b], "step " + std::to_string(s) +
                                                      " references an undefined or released slot");
    GRAFT_REQUIRE(!defined[st.dst] && st.dst >= ninputs, "step destination already defined");
    GRAFT_REQUIRE(st.a != st.b, "step contracts a slot with itself");
    GRAFT_REQUIRE(st.n_oindA >= 0 && st.n_cindA >= 0 && st.n_oindB >= 0 && st.n_cindB >= 0 &&
                      st.n_oindA <= GRAFT_MAX_RANK && st.n_oindB <= GRAFT_MAX_RANK &&
                      st.n_cindA <= GRAFT_MAX_RANK && st.n_cindB <= GRAFT_MAX_RANK &&
                      st.n_out_codomain >= 0 && st.n_out_domain >= 0 &&
                      st.n_out_codomain + st.n_out_domain <= GRAFT_MAX_RANK,
                  "step leg counts out of range");
    const TStruct* fixed = (st.dst == output_slot && have_out) ? &outfixed : nullptr;
    TStruct C;
    // hole rectangles are needed only for an operand a fused step can hand
    // over (fused, every result element mapped: compiled before this step)
    auto handable = [&](int slot) {
      if (slot < ninputs || slot_producer[slot] < 0) return false;
      const StepExec& e = p.steps[slot_producer[slot]];
      return e.fused && e.c_complete;
    };
    if (cp) {
      if (cp->npairs[s] < 0) err = "negative pair count";
      if (cp->npairs[s] < 0 || build_struct(cp->outs[s], C, err, true) ||
          compile_coupled_step(st, p.slots[st.a], p.slots[st.b], C, cp->pairs + 3 * pair_off,
                               cp->coefs + pair_off, cp->npairs[s], p.steps[s], err)) {
        set_error("step " + std::to_string(s) + ": " + err);
        return GRAFT_E_INVALID;
      }
      pair_off += cp->npairs[s];
    } else if (compile_step(st, p.slots[st.a], p.slots[st.b], fixed, C, p.steps[s], err,
                            ctx->sector_fuse, handable(st.a), handable(st.b))) {
      set_error("step " + std::to_string(s) + ": " + err);
      return GRAFT_E_INVALID;
    }
    p.slots[st.dst] = std::move(C);
    slot_producer[st.dst] = s;
    defined[st.dst] = 1;
    born[st.dst] = s;
    if (st.a >= ninputs) dies[st.a] = s;
    if (st.b >= ninputs) dies[st.b] = s;
    defined[st.a] = (st.a < ninputs);  // inputs stay valid; intermediates are consumed
    defined[st.b] = (st.b < ninputs);
    StepExec& ex = p.steps[s];
    ex.config = default_config(ex);
    assign_tiles(ex);
    if (ex.swork.size() >= (1u << 26)) {
      set_error("step " + std::to_string(s) + ": too many streaming work entries");
      return GRAFT_E_UNSUPPORTED;
    }
    p.macs += ex.macs;
    if (timing)
      std::fprintf(stderr, "plan step %d: %.2f ms, blocks A %zu B %zu C %zu, items %zu, fused %d, "
                 "stream %d lean %d (rank %d) %s\n", s,
                 std::chrono::duration<double, std::milli>(now() - t_step).count(),
                 p.slots[st.a].blocks.size(), p.slots[st.b].blocks.size(),
                 p.slots[st.dst].blocks.size(), ex.items.size(), (int)ex.fused, ex.bstream,
                 (int)ex.lean, ex.lean_rank, ex.lean_why);
  }
  GRAFT_REQUIRE(defined[output_slot], "output slot never produced");
  for (int s = ninputs; s < nslots; ++s)
    GRAFT_REQUIRE(s == output_slot || dies[s] >= 0, "an intermediate is never consumed");
  // Handoffs: a fused step whose result feeds a fused step as an operand
  // stores it in that step's fused layout (its slot grows to the fused
  // matrix when the component matrices hold hole rectangles).
  std::vector<int> producer(nslots, -1);
  std::vector<int64_t> need(nslots, 0);
  for (int s = 0; s < nsteps; ++s) producer[steps[s].dst] = s;
  for (int t = 0; t < nsteps; ++t) {
    StepExec& et = p.steps[t];
    if (!et.fused) {
      et.holes_a.clear();
      et.holes_b.clear();
      continue;
    }
    for (int side = 0; side < 2; ++side) {
      const int X = side == 0 ? steps[t].a : steps[t].b;
      const int s = (X >= ninputs && X != output_slot) ? producer[X] : -1;
      StepExec* es = s >= 0 ? &p.steps[s] : nullptr;
      if (!es || !es->fused || !es->c_complete || es->ho_step >= 0) {
        (side == 0 ? et.holes_a : et.holes_b).clear();
        continue;
      }
      es->ho_step = t;
      es->ho_side = side;
      if (side == 0) {
        et.ho_in_a = true;
        et.ho_from_a = s;
        need[X] = std::max(need[X], et.fa);
      } else {
        et.ho_in_b = true;
        et.ho_from_b = s;
        need[X] = std::max(need[X], et.fb);
      }
    }
  }
  // Split-word handoffs: a dense streaming step feeding a tiled GEMM step's A
  // operand (the TTNO absorb before the environment absorb).
  for (int t = 0; t < nsteps; ++t) {
    StepExec& et = p.steps[t];
    const int X = steps[t].a;
    const int s = (X >= ninputs && X != output_slot) ? producer[X] : -1;
    if (s < 0 || et.fused || p.steps[s].fused || p.steps[s].skinny == 0 ||
        p.steps[s].items.size() != 1)
      continue;
    p.steps[s].pl_to = t;
    et.pl_from = s;
    et.bplane_bytes = (uint64_t)p.slots[steps[t].b].total * 16;
  }
  // Workspace: first-fit over liveness intervals [born, dies].
  p.ws_offset.assign(nslots, -1);
  struct Live { int64_t off, size; int b, d; };
  std::vector<Live> placed;
  uint64_t peak = 0;
  for (int s = ninputs; s < nslots; ++s) {
    if (s == output_slot) continue;
    const int64_t size = ((std::max(p.slots[s].total, need[s]) * 16 + 255) / 256) * 256;
    const int b = born[s], d = dies[s];
    std::vector<std::pair<int64_t, int64_t>> busy;
    for (auto& l : placed)
      if (!(l.d < b || d < l.b)) busy.push_back({l.off, l.off + l.size});
    std::sort(busy.begin(), busy.end());
    int64_t off = 0;
    for (auto& iv : busy) {
      if (off + size <= iv.first) break;
      off = std::max(off, iv.second);
    }
    placed.push_back({off, size, b, d});
    p.ws_offset[s] = off;
    peak = std::max<uint64_t>(peak, off + size);
  }
  p.ws_slots = peak;
  update_workspace_size(p);
  auto t_up = now();
  rc = upload_step_meta(p);
  if (rc) return rc;
  auto t_maps = now();
  rc = build_maps(ctx, p);
  if (rc) return rc;
  if (timing)
    std::fprintf(stderr, "plan create: %.2f ms total, %.2f ms upload (%.1f MB), %.2f ms maps (%.1f MB)\n",
                 std::chrono::duration<double, std::milli>(now() - t_start).count(),
                 std::chrono::duration<double, std::milli>(t_maps - t_up).count(), p.meta_bytes / 1e6,
                 std::chrono::duration<double, std::milli>(now() - t_maps).count(), p.map_bytes / 1e6);
  *out = plan.release();
  return GRAFT_OK;
}

int graft_plan_create(graft_ctx* ctx, const graft_pair_step* steps, int32_t nsteps,
                      int32_t nslots, int32_t output_slot, const graft_tensor_desc* input_descs,
                      int32_t ninputs, const graft_tensor_desc* out_desc, graft_plan** out) {
  return plan_create_impl(ctx, steps, nsteps, nslots, output_slot, input_descs, ninputs, out_desc,
                          nullptr, out);
}

int graft_plan_create_coupled(graft_ctx* ctx, const graft_pair_step* steps, int32_t nsteps,
                              int32_t nslots, int32_t output_slot,
                              const graft_tensor_desc* input_descs, int32_t ninputs,
                              const graft_tensor_desc* step_out_descs, const int64_t* npairs,
                              const int32_t* pairs, const double* coefs, graft_plan** out) {
  GRAFT_REQUIRE(steps != nullptr && nsteps >= 1, "plan without steps");
  GRAFT_REQUIRE(step_out_descs != nullptr && npairs != nullptr,
                "coupled plan without step result descriptors or pair counts");
  int64_t total = 0;
  for (int s = 0; s < nsteps; ++s) total += npairs[s] > 0 ? npairs[s] : 0;
  GRAFT_REQUIRE(total == 0 || (pairs != nullptr && coefs != nullptr),
                "coupled plan without block pairs or coefficients");
  const CoupledSpec cp{step_out_descs, npairs, pairs, coefs};
  return plan_create_impl(ctx, steps, nsteps, nslots, output_slot, input_descs, ninputs, nullptr,
                          &cp, out);
}

int graft_plan_destroy(graft_plan* plan) {
  if (!plan) return GRAFT_OK;
  if (plan->impl.d_meta || plan->impl.d_maps) {
    (void)hipSetDevice(plan->impl.device);
    if (plan->impl.d_meta) (void)hipFree(plan->impl.d_meta);
    if (plan->impl.d_maps) (void)hipFree(plan->impl.d_maps);
  }
  delete plan;
  return GRAFT_OK;
}

int graft_plan_info(graft_plan* plan, uint64_t* workspace_bytes, double* macs,
                    uint64_t* output_elements) {
  GRAFT_REQUIRE(plan != nullptr, "null plan");
  if (workspace_bytes) *workspace_bytes = plan->impl.ws_bytes;
  if (macs) *macs = plan->impl.macs;
  if (output_elements) *output_elements = (uint64_t)plan->impl.slots[plan->impl.output_slot].total;
  return GRAFT_OK;
}

int graft_plan_device_bytes(graft_plan* plan, uint64_t* bytes) {
  GRAFT_REQUIRE(plan != nullptr && bytes != nullptr, "null plan or output");
  *bytes = plan->impl.meta_bytes + plan->impl.map_bytes;
  return GRAFT_OK;
}

int graft_plan_step_info(graft_plan* plan, int32_t step, int64_t* mkn3, int64_t* nblocks,
                         double* macs, int32_t* tile_config) {
  GRAFT_REQUIRE(plan != nullptr, "null plan");
  GRAFT_REQUIRE(step >= 0 && step < (int32_t)plan->impl.steps.size(), "step out of range");
  const StepExec& ex = plan->impl.steps[step];
  if (mkn3) {
    mkn3[0] = ex.maxM;
    mkn3[1] = ex.maxK;
    mkn3[2] = ex.maxN;
  }
  if (nblocks) *nblocks = (int64_t)(ex.cslices > 1 ? ex.items_base.size() : ex.items.size());
  if (macs) *macs = ex.macs;
  if (tile_config) *tile_config = ex.config;
  return GRAFT_OK;
}

int graft_plan_step_launch(graft_plan* plan, int32_t step, int32_t* config, int32_t* orientation,
                           int32_t* ksplit) {
  GRAFT_REQUIRE(plan != nullptr, "null plan");
  GRAFT_REQUIRE(step >= 0 && step < (int32_t)plan->impl.steps.size(), "step out of range");
  const StepExec& ex = plan->impl.steps[step];
  if (config) *config = ex.config;
  if (orientation) *orientation = ex.tr ? 1 : 0;
  if (ksplit) *ksplit = ex.cslices > 1 ? ex.cslices : ex.ksplit;
  return GRAFT_OK;
}

int graft_plan_step_elements(graft_plan* plan, int32_t step, int64_t* abc3) {
  GRAFT_REQUIRE(plan != nullptr && abc3 != nullptr, "null argument");
  GRAFT_REQUIRE(step >= 0 && step < (int32_t)plan->impl.steps.size(), "step out of range");
  const StepExec& ex = plan->impl.steps[step];
  abc3[0] = plan->impl.slots[ex.a].total;
  abc3[1] = plan->impl.slots[ex.b].total;
  abc3[2] = plan->impl.slots[ex.dst].total;
  return GRAFT_OK;
}

int graft_plan_step_gemms(graft_plan* plan, int32_t step, int64_t* ngemms, int32_t* mnks) {
  GRAFT_REQUIRE(plan != nullptr && ngemms != nullptr, "null argument");
  GRAFT_REQUIRE(step >= 0 && step < (int32_t)plan->impl.steps.size(), "step out of range");
  const StepExec& ex = plan->impl.steps[step];
  if (mnks == nullptr) {
    *ngemms = (int64_t)ex.items.size();
    return GRAFT_OK;
  }
  GRAFT_REQUIRE(*ngemms >= (int64_t)ex.items.size(), "shape array too small");
  for (size_t i = 0; i < ex.items.size(); ++i) {
    const GemmItem& it = ex.items[i];
    int64_t k = 0;
    for (int s = it.seg_begin; s < it.seg_begin + it.seg_count; ++s) k += ex.segs[s].K;
    mnks[4 * i] = it.M;
    mnks[4 * i + 1] = it.N;
    mnks[4 * i + 2] = (int32_t)k;
    mnks[4 * i + 3] = it.seg_count;
  }
  *ngemms = (int64_t)ex.items.size();
  return GRAFT_OK;
}

int graft_plan_output_blocks(graft_plan* plan, int32_t* rank, int64_t* nblocks,
                             int32_t* block_sectors, int64_t* block_offsets) {
  GRAFT_REQUIRE(plan != nullptr && rank != nullptr && nblocks != nullptr, "null argument");
  const TStruct& t = plan->impl.slots[plan->impl.output_slot];
  *rank = t.rank;
  if (block_sectors == nullptr || block_offsets == nullptr) {
    *nblocks = (int64_t)t.blocks.size();
    return GRAFT_OK;
  }
  GRAFT_REQUIRE(*nblocks >= (int64_t)t.blocks.size(), "block arrays too small");
  for (size_t b = 0; b < t.blocks.size(); ++b) {
    for (int l = 0; l < t.rank; ++l) block_sectors[b * t.rank + l] = t.blocks[b].sec[l];
    block_offsets[b] = t.blocks[b].offset;
  }
  *nblocks = (int64_t)t.blocks.size();
  return GRAFT_OK;
}

int graft_num_tile_configs(void) { return kNumTileConfigs + 2; }

int graft_plan_set_tile_config(graft_plan* plan, int32_t step, int32_t config) {
  GRAFT_REQUIRE(plan != nullptr, "null plan");
  GRAFT_REQUIRE(step >= 0 && step < (int32_t)plan->impl.steps.size(), "step out of range");
  GRAFT_REQUIRE(config >= -1 && config <= kBlockStreamConfig, "tile configuration out of range");
  StepExec& ex = plan->impl.steps[step];
  if (config == kSkinnyConfig && ex.skinny == 0) {
    set_error("the streaming kernel does not apply to this step");
    return GRAFT_E_UNSUPPORTED;
  }
  if (config == kBlockStreamConfig && ex.bstream == 0) {
    set_error("the per-block streaming kernel does not apply to this step");
    return GRAFT_E_UNSUPPORTED;
  }
  ex.pinned = config;
  ex.config = config >= 0 ? config : default_config(ex);
  assign_tiles(ex);
  update_workspace_size(plan->impl);
  return upload_step_meta(plan->impl);
}

int graft_plan_set_orientation(graft_plan* plan, int32_t step, int32_t orientation) {
  GRAFT_REQUIRE(plan != nullptr, "null plan");
  GRAFT_REQUIRE(step >= 0 && step < (int32_t)plan->impl.steps.size(), "step out of range");
  GRAFT_REQUIRE(orientation >= -1 && orientation <= 1, "orientation must be -1, 0 or 1");
  StepExec& ex = plan->impl.steps[step];
  ex.tr_pinned = orientation;
  ex.tr = orientation < 0 ? ex.tr_auto : orientation == 1;
  return GRAFT_OK;
}

int graft_plan_execute(graft_ctx* ctx, graft_plan* plan, const void* const* inputs, void* output,
                       int32_t precision) {
  GRAFT_REQUIRE(ctx != nullptr && plan != nullptr, "null argument");
  return execute_plan(ctx, plan->impl, inputs, output, precision);
}

int graft_plan_execute_timed(graft_ctx* ctx, graft_plan* plan, const void* const* inputs,
                             void* output, int32_t precision, float* step_ms) {
  GRAFT_REQUIRE(ctx != nullptr && plan != nullptr && step_ms != nullptr, "null argument");
  graft_plan_impl& p = plan->impl;
  GRAFT_REQUIRE(inputs != nullptr && output != nullptr, "null inputs/output");
  int rc = set_device(ctx);
  if (rc) return rc;
  rc = ensure_workspace(ctx, p.ws_bytes);
  if (rc) return rc;
  const size_t ns = p.steps.size();
  std::vector<hipEvent_t> ev(ns + 1);
  for (auto& e : ev) GRAFT_CHECK_HIP(hipEventCreate(&e));
  GRAFT_CHECK_HIP(hipEventRecord(ev[0], ctx->stream));
  for (size_t s = 0; s < ns && !rc; ++s) {
    rc = run_step(ctx, p, p.steps[s], inputs, output, precision, (char*)ctx->ws);
    if (!rc) GRAFT_CHECK_HIP(hipEventRecord(ev[s + 1], ctx->stream));
  }
  if (!rc) {
    GRAFT_CHECK_HIP(hipEventSynchronize(ev[ns]));
    for (size_t s = 0; s < ns; ++s) GRAFT_CHECK_HIP(hipEventElapsedTime(&step_ms[s], ev[s], ev[s + 1]));
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int graft_plan_autotune(graft_ctx* ctx, graft_plan* plan, const void* const* inputs, void* output,
                        int32_t precision, int32_t repeats) {
  GRAFT_REQUIRE(ctx != nullptr && plan != nullptr, "null argument");
  graft_plan_impl& p = plan->impl;
  int rc = execute_plan(ctx, p, inputs, output, precision);  // fills every slot once
  if (rc) return rc;
  if (repeats < 1) repeats = 1;
  static const bool log = std::getenv("GRAFT_AUTOTUNE_LOG") != nullptr;
  hipEvent_t e0, e1;
  GRAFT_CHECK_HIP(hipEventCreate(&e0));
  GRAFT_CHECK_HIP(hipEventCreate(&e1));
  int step_index = -1;
  for (auto& ex : p.steps) {
    ++step_index;
    if (ex.pinned >= 0 && ex.tr_pinned >= 0) continue;
    int best = ex.config;
    bool best_tr = ex.tr;
    float best_ms = 1e30f;
    for (int c = 0; c <= kBlockStreamConfig && !rc; ++c) {
      if (ex.pinned >= 0 && c != ex.pinned) continue;
      if (c == kSkinnyConfig && ex.skinny == 0) continue;
      if (c == kBlockStreamConfig && ex.bstream == 0) continue;
      ex.config = c;
      assign_tiles(ex);
      update_workspace_size(p);
      rc = ensure_workspace(ctx, p.ws_bytes);
      if (rc) break;
      rc = upload_step_meta(p);
      if (rc) break;
      for (int o = 0; o < 2 && !rc; ++o) {
        if (ex.tr_pinned >= 0 && o != ex.tr_pinned) continue;
        if ((c == kSkinnyConfig || c == kBlockStreamConfig) && o == 1) continue;
        ex.tr = (o == 1);
        // warm-up, then time
        rc = run_step(ctx, p, ex, inputs, output, precision, (char*)ctx->ws);
        if (rc) break;
        GRAFT_CHECK_HIP(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < repeats && !rc; ++r)
          rc = run_step(ctx, p, ex, inputs, output, precision, (char*)ctx->ws);
        GRAFT_CHECK_HIP(hipEventRecord(e1, ctx->stream));
        GRAFT_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        GRAFT_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (log)
          std::fprintf(stderr, "autotune step %d m=%lld n=%lld k=%lld blocks=%zu config %d tr %d: %.4f ms\n",
                       step_index, (long long)ex.maxM, (long long)ex.maxN, (long long)ex.maxK, ex.items.size(), c,
                       (int)ex.tr, ms / repeats);
        if (!rc && ms < best_ms) {
          best_ms = ms;
          best = c;
          best_tr = ex.tr;
        }
      }
    }
    if (rc) break;
    ex.config = best;
    ex.tr = best_tr;
    assign_tiles(ex);
    update_workspace_size(p);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  rc = upload_step_meta(p);
  if (rc) return rc;
  // Re-run so the output holds the plan's result.
  return execute_plan(ctx, p, inputs, output, precision);
}

int graft_heff_apply(graft_ctx* ctx, graft_plan* plan, const void* x, const void* const* statics,
                     void* y, int32_t precision) {
  GRAFT_REQUIRE(ctx != nullptr && plan != nullptr, "null argument");
  const int n = plan->impl.ninputs;
  GRAFT_REQUIRE(n >= 1 && x != nullptr && (n == 1 || statics != nullptr), "null operand");
  std::vector<const void*> inputs(n);
  inputs[0] = x;
  for (int i = 1; i < n; ++i) inputs[i] = statics[i - 1];
  return execute_plan(ctx, plan->impl, inputs.data(), y, precision);
}

int graft_heff_apply_noalloc(graft_ctx* ctx, graft_plan* plan, const void* x,
                             const void* const* statics, void* y, int32_t precision) {
  GRAFT_REQUIRE(ctx != nullptr && plan != nullptr, "null argument");
  graft_plan_impl& p = plan->impl;
  const int n = p.ninputs;
  GRAFT_REQUIRE(n >= 1 && x != nullptr && (n == 1 || statics != nullptr) && y != nullptr,
                "null operand");
  GRAFT_REQUIRE(n <= kNoallocMaxInputs, "too many plan inputs for the allocation-free apply");
  GRAFT_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices, "device index out of range");
  if (p.ws_bytes > kNoallocScratchBytes) {
    set_error("the plan needs " + std::to_string(p.ws_bytes) +
              " workspace bytes; the allocation-free apply has a static scratch of " +
              std::to_string(kNoallocScratchBytes) + " bytes");
    return GRAFT_E_NOMEM;
  }
  char* ws = (char*)g_scratch_ptr[ctx->device];
  GRAFT_REQUIRE(ws != nullptr, "the static scratch of this device is not set up");
  int rc = set_device(ctx);
  if (rc) return rc;
  const void* inputs[kNoallocMaxInputs];
  inputs[0] = x;
  for (int i = 1; i < n; ++i) inputs[i] = statics[i - 1];
  for (auto& ex : p.steps) {
    rc = run_step(ctx, p, ex, inputs, y, precision, ws);
    if (rc) return rc;
  }
  return GRAFT_OK;
}

int graft_env_update(graft_ctx* ctx, graft_plan* plan, const void* const* operands,
                     void* env_out) {
  GRAFT_REQUIRE(ctx != nullptr && plan != nullptr, "null argument");
  return execute_plan(ctx, plan->impl, operands, env_out, GRAFT_PREC_FP64);
}

int graft_contract_pair(graft_ctx* ctx, const graft_pair_step* step,
                        const graft_tensor_desc* descA, const void* A,
                        const graft_tensor_desc* descB, const void* B, void* C) {
  GRAFT_REQUIRE(ctx && step && descA && descB && A && B && C, "null argument");
  graft_pair_step st = *step;
  st.a = 0;
  st.b = 1;
  st.dst = 2;
  graft_tensor_desc descs[2] = {*descA, *descB};
  graft_plan* plan = nullptr;
  int rc = graft_plan_create(ctx, &st, 1, 3, 2, descs, 2, nullptr, &plan);
  if (rc) return rc;
  const void* ins[2] = {A, B};
  rc = graft_plan_execute(ctx, plan, ins, C, GRAFT_PREC_FP64);
  if (rc == GRAFT_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      set_error(std::string("HIP error: ") + hipGetErrorString(e));
      rc = GRAFT_E_HIP;
    }
  }
  graft_plan_destroy(plan);
  return rc;
}

int graft_vec_dotc(graft_ctx* ctx, int64_t n, const void* x, const void* y, double* out2) {
  GRAFT_REQUIRE(ctx && x && y && out2 && n >= 0, "bad argument");
  int rc = set_device(ctx);
  if (rc) return rc;
  GRAFT_CHECK_HIP(launch_dotc_partials(n, x, y, ctx->d_partials, kReduceBlocks, ctx->stream));
  GRAFT_CHECK_HIP(launch_reduce_partials(ctx->d_partials, kReduceBlocks, 1, ctx->d_result, ctx->stream));
  GRAFT_CHECK_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, 2 * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
  GRAFT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  out2[0] = ctx->h_result[0];
  out2[1] = ctx->h_result[1];
  return GRAFT_OK;
}

int graft_vec_nrm2(graft_ctx* ctx, int64_t n, const void* x, double* out) {
  double d[2];
  int rc = graft_vec_dotc(ctx, n, x, x, d);
  if (rc) return rc;
  *out = std::sqrt(std::max(d[0], 0.0));
  return GRAFT_OK;
}

int graft_vec_axpby(graft_ctx* ctx, int64_t n, const double* alpha, const void* x,
                    const double* beta, void* y) {
  GRAFT_REQUIRE(ctx && alpha && beta && x && y && n >= 0, "bad argument");
  int rc = set_device(ctx);
  if (rc) return rc;
  GRAFT_CHECK_HIP(launch_axpby(n, alpha[0], alpha[1], x, beta[0], beta[1], y, ctx->stream));
  return GRAFT_OK;
}

static PtrList make_ptrs(int k, const void* const* V) {
  PtrList L;
  for (int j = 0; j < kMaxMultiVec; ++j) L.p[j] = j < k ? (const double2*)V[j] : nullptr;
  return L;
}

int graft_vec_multi_dotc(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                         const void* y, double* out2k) {
  GRAFT_REQUIRE(ctx && V && y && out2k && n >= 0, "bad argument");
  GRAFT_REQUIRE(k >= 1 && k <= kMaxMultiVec, "vector count out of range");
  int rc = set_device(ctx);
  if (rc) return rc;
  GRAFT_CHECK_HIP(launch_multi_dotc_partials(n, k, make_ptrs(k, V), y, ctx->d_partials,
                                             kReduceBlocks, ctx->stream));
  GRAFT_CHECK_HIP(launch_reduce_partials(ctx->d_partials, kReduceBlocks, k, ctx->d_result,
                                         ctx->stream));
  GRAFT_CHECK_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, 2 * k * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
  GRAFT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(out2k, ctx->h_result, 2 * k * sizeof(double));
  return GRAFT_OK;
}

int graft_vec_multi_axpy(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                         const double* c2k, void* y) {
  GRAFT_REQUIRE(ctx && V && c2k && y && n >= 0, "bad argument");
  GRAFT_REQUIRE(k >= 1 && k <= kMaxMultiVec, "vector count out of range");
  int rc = set_device(ctx);
  if (rc) return rc;
  // The coefficient table is reused by the next call, so the upload is
  // ordered after earlier kernels on the stream.
  GRAFT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  GRAFT_CHECK_HIP(hipMemcpy(ctx->d_coef, c2k, sizeof(double) * 2 * k, hipMemcpyHostToDevice));
  GRAFT_CHECK_HIP(launch_multi_axpy(n, k, make_ptrs(k, V), ctx->d_coef, 1.0, y, ctx->stream));
  return GRAFT_OK;
}

int graft_vec_multi_dotc_dev(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                             const void* y, double* d_out2k) {
  GRAFT_REQUIRE(ctx && V && y && d_out2k && n >= 0, "bad argument");
  GRAFT_REQUIRE(k >= 1 && k <= kMaxMultiVec, "vector count out of range");
  GRAFT_CHECK_HIP(launch_multi_dotc_partials(n, k, make_ptrs(k, V), y, ctx->d_partials,
                                             kReduceBlocks, ctx->stream));
  GRAFT_CHECK_HIP(launch_reduce_partials(ctx->d_partials, kReduceBlocks, k, d_out2k, ctx->stream));
  return GRAFT_OK;
}

int graft_vec_multi_axpy_dev(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                             const double* d_coef2k, double sign, void* y) {
  GRAFT_REQUIRE(ctx && V && d_coef2k && y && n >= 0, "bad argument");
  GRAFT_REQUIRE(k >= 1 && k <= kMaxMultiVec, "vector count out of range");
  GRAFT_CHECK_HIP(launch_multi_axpy(n, k, make_ptrs(k, V), d_coef2k, sign, y, ctx->stream));
  return GRAFT_OK;
}

int graft_vec_orthonormalize_dev(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                                 void* w, double eta, double* d_c1, double* d_c2, double* d_beta2) {
  GRAFT_REQUIRE(ctx && V && w && d_c1 && d_c2 && d_beta2 && n >= 0, "bad argument");
  GRAFT_REQUIRE(k >= 1 && k <= kMaxMultiVec, "vector count out of range");
  GRAFT_REQUIRE(eta > 0.0 && eta < 1.0, "eta must lie in (0, 1)");
  const PtrList P = make_ptrs(k, V);
  double* gate = d_beta2 + 1;
  hipStream_t s = ctx->stream;
  // first pass
  GRAFT_CHECK_HIP(launch_multi_dotc_partials(n, k, P, w, ctx->d_partials, kReduceBlocks, s));
  GRAFT_CHECK_HIP(launch_reduce_partials(ctx->d_partials, kReduceBlocks, k, d_c1, s));
  GRAFT_CHECK_HIP(launch_multi_axpy(n, k, P, d_c1, -1.0, w, s));
  // ||w_1|| and the refinement decision, on the device
  GRAFT_CHECK_HIP(launch_dotc_partials(n, w, w, ctx->d_partials, kReduceBlocks, s));
  GRAFT_CHECK_HIP(launch_reduce_partials(ctx->d_partials, kReduceBlocks, 1, ctx->d_result, s));
  GRAFT_CHECK_HIP(launch_reorth_gate(d_c1, k, ctx->d_result, eta, d_beta2, gate, s));
  // second pass (skipped inside the kernels when the gate is off)
  GRAFT_CHECK_HIP(launch_multi_dotc_partials(n, k, P, w, ctx->d_partials, kReduceBlocks, s, gate));
  GRAFT_CHECK_HIP(launch_reduce_partials(ctx->d_partials, kReduceBlocks, k, d_c2, s));
  GRAFT_CHECK_HIP(launch_multi_axpy(n, k, P, d_c2, -1.0, w, s, gate));
  GRAFT_CHECK_HIP(launch_dotc_partials(n, w, w, ctx->d_partials, kReduceBlocks, s, gate));
  GRAFT_CHECK_HIP(launch_reduce_partials(ctx->d_partials, kReduceBlocks, 1, ctx->d_result, s));
  GRAFT_CHECK_HIP(launch_reorth_finish(ctx->d_result, gate, d_beta2, s));
  GRAFT_CHECK_HIP(launch_scale_inv(n, d_beta2, w, s));
  return GRAFT_OK;
}

int graft_vec_nrm2_dev(graft_ctx* ctx, int64_t n, const void* x, double* d_out) {
  GRAFT_REQUIRE(ctx && x && d_out && n >= 0, "bad argument");
  GRAFT_CHECK_HIP(launch_dotc_partials(n, x, x, ctx->d_partials, kReduceBlocks, ctx->stream));
  GRAFT_CHECK_HIP(launch_reduce_partials(ctx->d_partials, kReduceBlocks, 1, ctx->d_result,
                                         ctx->stream));
  GRAFT_CHECK_HIP(launch_sqrt_real(ctx->d_result, d_out, ctx->stream));
  return GRAFT_OK;
}

int graft_vec_scale_inv_dev(graft_ctx* ctx, int64_t n, const double* d_beta, void* x) {
  GRAFT_REQUIRE(ctx && d_beta && x && n >= 0, "bad argument");
  GRAFT_CHECK_HIP(launch_scale_inv(n, d_beta, x, ctx->stream));
  return GRAFT_OK;
}

int graft_vec_block_combine(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                            int32_t p, const double* coef, void* const* O) {
  GRAFT_REQUIRE(ctx && V && coef && O && n >= 0, "bad argument");
  GRAFT_REQUIRE(k >= 1 && k <= kMaxMultiVec, "input vector count out of range");
  GRAFT_REQUIRE(p >= 1 && p <= kMaxCombine, "output vector count out of range");
  GRAFT_REQUIRE(!ctx->capturing, "block combine uploads coefficients and cannot be captured");
  int rc = set_device(ctx);
  if (rc) return rc;
  // stream-ordered upload: kernels queued before this call finish reading
  // the coefficient buffer before it is overwritten
  GRAFT_CHECK_HIP(hipMemcpyAsync(ctx->d_comb, coef, sizeof(double) * 2 * k * p,
                                 hipMemcpyHostToDevice, ctx->stream));
  GRAFT_CHECK_HIP(launch_block_combine(n, k, make_ptrs(k, V), p, make_ptrs(p, (const void* const*)O),
                                       ctx->d_comb, ctx->stream));
  return GRAFT_OK;
}

int graft_graph_capture_begin(graft_ctx* ctx) {
  GRAFT_REQUIRE(ctx != nullptr, "null context");
  GRAFT_REQUIRE(ctx->stream != nullptr, "graph capture needs a non-default stream");
  GRAFT_REQUIRE(!ctx->capturing, "a capture is already open on this context");
  GRAFT_CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  return GRAFT_OK;
}

int graft_graph_capture_end(graft_ctx* ctx, graft_graph** out) {
  GRAFT_REQUIRE(ctx != nullptr && out != nullptr, "null argument");
  GRAFT_REQUIRE(ctx->capturing, "no capture is open on this context");
  ctx->capturing = false;
  *out = nullptr;
  hipGraph_t g = nullptr;
  GRAFT_CHECK_HIP(hipStreamEndCapture(ctx->stream, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    set_error(std::string("graph instantiation failed: ") + hipGetErrorString(e));
    return GRAFT_E_HIP;
  }
  graft_graph* gr = new graft_graph;
  gr->exec = ex;
  gr->ctx = ctx;
  gr->ws_generation = ctx->ws_generation;
  *out = gr;
  return GRAFT_OK;
}

int graft_graph_launch(graft_ctx* ctx, graft_graph* graph) {
  GRAFT_REQUIRE(ctx != nullptr && graph != nullptr, "null argument");
  GRAFT_REQUIRE(graph->ctx == ctx, "the graph was captured on another context");
  GRAFT_REQUIRE(graph->ws_generation == ctx->ws_generation,
                "the context's workspace moved since the graph was captured: capture it again");
  GRAFT_CHECK_HIP(hipGraphLaunch(graph->exec, ctx->stream));
  return GRAFT_OK;
}

int graft_graph_destroy(graft_graph* graph) {
  if (graph) {
    (void)hipGraphExecDestroy(graph->exec);
    delete graph;
  }
  return GRAFT_OK;
}

}  // extern "C"
