/*
 * graft_heff.h — C ABI of the MI355X effective-Hamiltonian engine.
 *
 * The engine executes the pairwise contraction plans that the reference's
 * planner produces for eff_h1 / eff_h2 / eff_h0 applies and for the
 * environment update, on device-resident complex tensors, with hand-written
 * CDNA4 (gfx950) MFMA kernels.  Every entry point takes plain pointers and
 * sizes, returns an int status (GRAFT_OK == 0) and records a message that
 * graft_last_error() returns (per calling thread).
 *
 * Tensor storage follows TensorKit's block layout for a TensorMap with
 * trivial sectors: one dense block, flat legs (codomain legs, then domain
 * legs), first leg fastest, complex128 interleaved (re, im).  Abelian
 * sector tensors are a list of strided dense blocks addressed by one sector
 * index per leg (graft_tensor_desc below), which covers the layout TensorKit
 * uses for each coupled-sector block of a U(1) x U(1) TensorMap.
 *
 * Reference interfaces replaced (paths relative to the Graft.jl repo):
 *   graft_plan_create   <- Planning.PairStep / ContractionPlan
 *                          lib/GraftPlanning/src/Planning/types.jl:54-65, 94-115
 *   graft_plan_create_coupled <- the same plans over SU(2) TensorMaps, whose
 *                          pairwise steps TensorKit runs as fusion-tree
 *                          recoupling + one GEMM per coupled sector inside
 *                          Backend.contract_pair! (Backend.jl:209-222), with
 *                          the planner's dense selection for non-unique
 *                          fusion (lib/GraftPlanning/src/Planning/planner.jl:566-583)
 *   graft_plan_execute  <- Planning.execute / _execute_workspace
 *                          lib/GraftPlanning/src/Planning/executor.jl:16-45, 99-143
 *   graft_heff_apply    <- (::EffectiveMap)(x) / (::WorkspaceMap)(x)
 *                          lib/GraftPlanning/src/Planning/types.jl:440-445, 565-574
 *   graft_env_update    <- Contractions._build_env_value -> Planning.execute
 *                          lib/GraftContractions/src/Contractions/envcache.jl:1127-1132
 *   graft_contract_pair <- Backend.contract_pair / contract_pair!
 *                          lib/GraftFoundation/src/Backend/Backend.jl:166-222
 *   graft_vec_*         <- the vector operations KrylovKit performs inside
 *                          eigsolve / exponentiate on the map's output
 *                          (lib/GraftGroundState/src/GroundState/GroundState.jl:25-39,
 *                           lib/GraftEvolution/src/Evolution/Evolution.jl:134-148)
 */
#ifndef GRAFT_HEFF_H
#define GRAFT_HEFF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRAFT_MAX_RANK 12

enum graft_status {
  GRAFT_OK = 0,
  GRAFT_E_INVALID = 1,      /* malformed plan, descriptor or argument        */
  GRAFT_E_NOMEM = 2,        /* device allocation or byte-cap admission failed */
  GRAFT_E_HIP = 3,          /* a HIP runtime call failed                     */
  GRAFT_E_UNSUPPORTED = 4   /* valid request outside what this build supports */
};

/* Arithmetic of the complex products (the Krylov algebra stays FP64). */
enum graft_precision {
  GRAFT_PREC_FP64 = 0,      /* complex FP64 on v_mfma_f64_16x16x4f64          */
  GRAFT_PREC_FP32 = 1,      /* operands rounded to FP32, FP32 MFMA, FP32 acc.  */
  GRAFT_PREC_BF16X3 = 2     /* FP32 split into bf16 hi+lo, 3 bf16 MFMA products,
                               FP32 accumulation                               */
};

typedef struct graft_ctx graft_ctx;
typedef struct graft_plan graft_plan;
typedef struct graft_graph graft_graph;

/*
 * One binary contraction, the reference's PairStep with 0-based indices:
 * C = contract(A = slot[a] (conj if conjA), B = slot[b] (conj if conjB)),
 * contracting A legs cindA[i] with B legs cindB[i]; the result's flat legs
 * are out[0..n_out-1] indexing the concatenation (A legs oindA..., B legs
 * oindB...).  n_out_codomain only records the TensorMap partition.
 */
typedef struct graft_pair_step {
  int32_t a, b, dst;
  int32_t n_oindA, n_cindA, n_oindB, n_cindB;
  int32_t oindA[GRAFT_MAX_RANK];
  int32_t cindA[GRAFT_MAX_RANK];
  int32_t oindB[GRAFT_MAX_RANK];
  int32_t cindB[GRAFT_MAX_RANK];
  int32_t conjA, conjB;
  int32_t n_out_codomain, n_out_domain;
  int32_t out[GRAFT_MAX_RANK];
} graft_pair_step;

/*
 * Structure of one tensor.  Dense tensors: rank and dims, all pointers NULL,
 * nblocks = 1.  Sector tensors: nsectors[l] sectors on leg l with sizes
 * sector_dims (concatenated leg by leg); nblocks blocks, block b covering
 * sector indices block_sectors[b*rank + l]; its data starts at element
 * block_offsets[b] of the buffer with per-leg element strides
 * block_strides[b*rank + l] (NULL = packed, first leg fastest, blocks in
 * order).  Sector indices are matched by position between contracted legs;
 * the caller supplies charge-conserving block lists.
 */
typedef struct graft_tensor_desc {
  int32_t rank;
  int64_t dims[GRAFT_MAX_RANK];
  int32_t nsectors[GRAFT_MAX_RANK];
  const int64_t* sector_dims;
  int64_t nblocks;
  const int32_t* block_sectors;
  const int64_t* block_offsets;
  const int64_t* block_strides;
} graft_tensor_desc;

/* ---- context: one device, one stream, one workspace arena ------------- */
int graft_ctx_create(int device, void* hip_stream, graft_ctx** out);
int graft_ctx_destroy(graft_ctx* ctx);
int graft_ctx_set_stream(graft_ctx* ctx, void* hip_stream);
/* Hard cap on workspace bytes (0 = no cap); exceeding it fails with
 * GRAFT_E_NOMEM before any launch (mirrors memory_cap_bytes). */
int graft_ctx_set_workspace_cap(graft_ctx* ctx, uint64_t bytes);
int graft_ctx_workspace_bytes(graft_ctx* ctx, uint64_t* bytes);
int graft_ctx_release_workspace(graft_ctx* ctx);
/* Grow the workspace arena to at least `bytes` now (plans executed inside a
 * graph capture must find their workspace allocated).  The reference sizes
 * its PlanWorkspace up front the same way (types.jl:539-554). */
int graft_ctx_reserve_workspace(graft_ctx* ctx, uint64_t bytes);
/* How plans created afterwards run block-sparse steps: 0 one GEMM item per
 * output block; 1 (default) automatic: per-block streaming (blocks read in
 * place, FP64) for steps whose smaller operand adds at most 16 open elements
 * to each output block, else per-charge fused matrices (gather, one grouped
 * dense GEMM over the charge components, scatter) for steps made of many
 * small blocks; 2 fused for every multi-block step; 3 per-block streaming for
 * every multi-block step.  TensorKit multiplies per coupled sector the same
 * way (BlockSparse mul!, as in Backend.jl:166-222). */
int graft_ctx_set_sector_fusion(graft_ctx* ctx, int32_t mode);
/* How fused FP64 MFMA sector steps reach their operands: 1 (default) the GEMM
 * reads the blocks and writes the result blocks through per-element maps in
 * its tile loads and stores (one launch per step); 0 gather the per-charge
 * matrices, multiply, scatter (three more launches).  Same results.  No
 * reference counterpart (TensorKit multiplies block by block). */
int graft_ctx_set_fused_map_loads(graft_ctx* ctx, int32_t on);
/* Which kernel runs the per-block streaming sector steps: 1 (default) the
 * lean kernel where it applies (small operand of at most 4096 elements,
 * staged in LDS per workgroup; at most 16 small-side outputs per output
 * block), 0 always the general per-block kernel.  Same arithmetic, same
 * summation order.  No reference counterpart (TensorKit multiplies block by
 * block). */
int graft_ctx_set_lean_stream(graft_ctx* ctx, int32_t on);
/* BF16X3 applies: 1 (default) a dense streaming step whose result is a
 * tiled GEMM step's A operand writes it as split-bf16 word pairs (8 bytes per
 * element) and the GEMM stages them, with its B operand split once per apply,
 * without converting on every tile load; 0 the GEMM converts FP64 operands
 * itself.  Same values.  No reference counterpart (the reference multiplies
 * in FP64 only). */
int graft_ctx_set_split_words(graft_ctx* ctx, int32_t on);
int graft_synchronize(graft_ctx* ctx);
const char* graft_last_error(void);
const char* graft_version(void);

/* ---- device block store ------------------------------------------------ */
int graft_malloc(graft_ctx* ctx, uint64_t bytes, void** dptr);
int graft_free(graft_ctx* ctx, void* dptr);
int graft_upload(graft_ctx* ctx, void* dst, const void* host_src, uint64_t bytes);
int graft_download(graft_ctx* ctx, void* host_dst, const void* src, uint64_t bytes);

/* ---- plans ------------------------------------------------------------- */
/* input_descs[0..ninputs-1] describe slots 0..ninputs-1 (slot 0 is the
 * Krylov vector x for effective maps); out_desc describes the result
 * (NULL = derive it from the plan).  Shapes are fixed at creation. */
int graft_plan_create(graft_ctx* ctx, const graft_pair_step* steps, int32_t nsteps,
                      int32_t nslots, int32_t output_slot,
                      const graft_tensor_desc* input_descs, int32_t ninputs,
                      const graft_tensor_desc* out_desc, graft_plan** out);
/*
 * A plan over non-abelian (SU(2)) block tensors.  Every tensor is a list of
 * degeneracy blocks (graft_tensor_desc; one block per sector tuple AND fusion
 * tree, so blocks may repeat a sector tuple and are addressed by position).
 * step_out_descs[s] lists the blocks of step s's result.  Step s multiplies
 * npairs[s] block pairs: pairs holds (C block, A block, B block) index
 * triples, the steps' lists concatenated, and coefs one real factor per pair,
 *   C_blk[c] = sum over its pairs of coef * contract(op(A_blk[a]), op(B_blk[b])),
 * the pairs of one output block summed in the order given; an output block
 * without pairs is zero.  The caller computes the factors (recoupling
 * coefficients of the blocks' fusion trees).  Steps run on the FP64 tile
 * kernels; other precisions return GRAFT_E_UNSUPPORTED at execution.
 */
int graft_plan_create_coupled(graft_ctx* ctx, const graft_pair_step* steps, int32_t nsteps,
                              int32_t nslots, int32_t output_slot,
                              const graft_tensor_desc* input_descs, int32_t ninputs,
                              const graft_tensor_desc* step_out_descs, const int64_t* npairs,
                              const int32_t* pairs, const double* coefs, graft_plan** out);
int graft_plan_destroy(graft_plan* plan);
/* Workspace bytes the plan needs, algorithmic complex MACs (sum m*k*n over
 * all GEMM blocks), and the number of output elements. */
int graft_plan_info(graft_plan* plan, uint64_t* workspace_bytes, double* macs,
                    uint64_t* output_elements);
/* Device bytes the plan itself holds (block tables and the per-element maps
 * of fused sector steps; the workspace arena belongs to the context).  The
 * reference's plans are host-only and carry no such figure; a plan cache
 * bounds its device memory with it. */
int graft_plan_device_bytes(graft_plan* plan, uint64_t* bytes);
/* Per-step GEMM profile: m, k, n of the largest block, block count and MACs. */
int graft_plan_step_info(graft_plan* plan, int32_t step, int64_t* mkn3, int64_t* nblocks,
                         double* macs, int32_t* tile_config);
/* Launch shape chosen for one step: tile configuration (numbered as in
 * graft_plan_set_tile_config), MFMA orientation, split-K factor (for a
 * coupled step: the chunks every output block's pair list is cut into). */
int graft_plan_step_launch(graft_plan* plan, int32_t step, int32_t* config, int32_t* orientation,
                           int32_t* ksplit);
/* Stored complex elements of one step's operands and result (A, B, C): the
 * step's algorithmic HBM bytes are 16 * (A + B + C).  No counterpart in the
 * reference (its pair_cost, Backend.jl:328-368, counts FLOPs only). */
int graft_plan_step_elements(graft_plan* plan, int32_t step, int64_t* abc3);
/* GEMM items of one step (output blocks, or per-charge components of a
 * fused sector step): call with mnks NULL to get their number, then with
 * 4 * ngemms int32 receiving (M, N, summed K, segments) per item.  No
 * reference counterpart: shape statistics for tile selection and tests. */
int graft_plan_step_gemms(graft_plan* plan, int32_t step, int64_t* ngemms, int32_t* mnks);
/* Block list of the plan's result (sector tensors): call with NULL arrays to
 * get rank and nblocks, then with arrays of nblocks*rank sector indices and
 * nblocks element offsets (packed, in this order). */
int graft_plan_output_blocks(graft_plan* plan, int32_t* rank, int64_t* nblocks,
                             int32_t* block_sectors, int64_t* block_offsets);
/* Pin the tile configuration of one step (-1 = heuristic).  Configurations
 * 0 .. n-3 are MFMA tile shapes, n-2 the streaming kernel of single-block
 * steps, n-1 the per-block streaming kernel of multi-block steps
 * (n = graft_num_tile_configs()); a kernel that does not apply to the step
 * returns GRAFT_E_UNSUPPORTED. */
int graft_plan_set_tile_config(graft_plan* plan, int32_t step, int32_t config);
int graft_num_tile_configs(void);
/* Pin the MFMA orientation of one step: 0 = rows on lanes 16-group,
 * 1 = transposed (B^T A^T), -1 = from the output's unit-stride leg. */
int graft_plan_set_orientation(graft_plan* plan, int32_t step, int32_t orientation);
/* Time every tile configuration and both MFMA orientations of every step on
 * the given operands and keep the fastest (the tile autotuner).  Operands
 * are left unchanged; the output holds the plan's result afterwards. */
int graft_plan_autotune(graft_ctx* ctx, graft_plan* plan, const void* const* inputs,
                        void* output, int32_t precision, int32_t repeats);
int graft_plan_execute(graft_ctx* ctx, graft_plan* plan, const void* const* inputs,
                       void* output, int32_t precision);
/* Same, with a HIP event pair on the context's stream around every step's
 * launch; step_ms[s] receives that step's duration (waits for completion). */
int graft_plan_execute_timed(graft_ctx* ctx, graft_plan* plan, const void* const* inputs,
                             void* output, int32_t precision, float* step_ms);

/* eff_h1 / eff_h2 / eff_h0 apply: inputs = (x, statics...). */
int graft_heff_apply(graft_ctx* ctx, graft_plan* plan, const void* x,
                     const void* const* statics, void* y, int32_t precision);
/* graft_heff_apply that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The plan's
 * intermediates, split-K partials and fused sector matrices live in a static
 * device array of 128 MiB per device instead of the context's workspace
 * arena; a plan needing more returns GRAFT_E_NOMEM.  Every call on one device
 * shares that array and leaves its contents behind, so calls that overlap on
 * one device (several streams or threads) race. */
int graft_heff_apply_noalloc(graft_ctx* ctx, graft_plan* plan, const void* x,
                             const void* const* statics, void* y, int32_t precision);
/* env(u -> v) from (ket, envs..., cap?, W, bra). */
int graft_env_update(graft_ctx* ctx, graft_plan* plan, const void* const* operands,
                     void* env_out);
/* One dense binary contraction (two_site_tensor, absorb_on_leg). */
int graft_contract_pair(graft_ctx* ctx, const graft_pair_step* step,
                        const graft_tensor_desc* descA, const void* A,
                        const graft_tensor_desc* descB, const void* B, void* C);

/* ---- Krylov vector operations on complex128 vectors of length n ------- */
int graft_vec_dotc(graft_ctx* ctx, int64_t n, const void* x, const void* y, double* out2);
int graft_vec_nrm2(graft_ctx* ctx, int64_t n, const void* x, double* out);
/* y = alpha*x + beta*y (alpha, beta complex: re, im pairs) */
int graft_vec_axpby(graft_ctx* ctx, int64_t n, const double* alpha, const void* x,
                    const double* beta, void* y);
/* out[j] = <V_j, y> for j < k (one pass over y) */
int graft_vec_multi_dotc(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                         const void* y, double* out2k);
/* y += sum_j c_j V_j (one pass over y) */
int graft_vec_multi_axpy(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                         const double* c2k, void* y);

/* Device-resident variants (results and coefficients stay in device memory,
 * nothing synchronizes): the building blocks of a captured Lanczos step.
 * They replace the same KrylovKit vector operations as above
 * (lib/GraftGroundState/src/GroundState/GroundState.jl:25-26,
 * lib/GraftEvolution/src/Evolution/Evolution.jl:134-135); the reference has no
 * device-scalar form: KrylovKit reads every inner product on the host. */
int graft_vec_multi_dotc_dev(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                             const void* y, double* d_out2k);
/* y += sign * sum_j c_j V_j with c read from device memory at run time */
int graft_vec_multi_axpy_dev(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                             const double* d_coef2k, double sign, void* y);
int graft_vec_nrm2_dev(graft_ctx* ctx, int64_t n, const void* x, double* d_out);
int graft_vec_scale_inv_dev(graft_ctx* ctx, int64_t n, const double* d_beta, void* x);
/* One Krylov orthonormalization of w against the orthonormal V_0..V_{k-1},
 * all on the device and capturable: a classical Gram-Schmidt pass (d_c1,
 * 2k doubles), a second pass (d_c2, zero when skipped) only when the first
 * removed more than 1 - eta of the norm -- the refinement rule of KrylovKit's
 * ModifiedGramSchmidtIR (eta = 1/sqrt(2)), the orthogonalizer behind the
 * eigsolve / exponentiate hooks (GroundState.jl:25-26, Evolution.jl:134-135)
 * -- then w /= ||w||.  d_beta2[0] = ||w|| before the scaling, d_beta2[1] = 1
 * when the second pass ran, else 0. */
int graft_vec_orthonormalize_dev(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                                 void* w, double eta, double* d_c1, double* d_c2, double* d_beta2);
/* O_j = sum_l coef[l + j*k] V_l for j < p (host coefficients, column-major
 * k x p complex, p <= 32).  Outputs may alias inputs, so a Krylov basis is
 * recombined in place at a thick restart.  Not capturable. */
int graft_vec_block_combine(graft_ctx* ctx, int64_t n, int32_t k, const void* const* V,
                            int32_t p, const double* coef, void* const* O);

/* ---- HIP graphs of work issued on the context's stream ---------------- */
/* No reference counterpart (the reference has no device): whole Lanczos
 * iterations are captured once and replayed at every restart.  Capture every
 * launch issued on the context's stream between begin and end (plans must
 * have their workspace already: graft_ctx_reserve_workspace or one run).
 * A graph refuses to launch once the context's workspace has moved. */
int graft_graph_capture_begin(graft_ctx* ctx);
int graft_graph_capture_end(graft_ctx* ctx, graft_graph** out);
int graft_graph_launch(graft_ctx* ctx, graft_graph* graph);
int graft_graph_destroy(graft_graph* graph);

#ifdef __cplusplus
}
#endif
#endif /* GRAFT_HEFF_H */
