/*
 * snicit.h -- C-ABI of the MI355X (gfx950) SNICIT sparse-DNN inference layer loop.
 *
 * Plain C types only (no HIP, no torch): every entry point takes host pointers and
 * sizes, returns 0 on success and a negative SNICIT_E* code on failure, and never
 * throws across the boundary.  snicit_last_error() gives the message of the last
 * failure on the calling thread.
 *
 * The reference has no library API; its in-process boundary is a set of header-only
 * classes whose constructor takes the network and whose infer() runs the layer loop:
 *   SNICIT_SDGC::SNICIT(weight_path, bias, N, L, threshold).infer(input, golden, num_inputs, batch)
 *        src/SDGC/SNICIT/SNICIT.hpp:43-56, 87-203
 *   SNICIT_SDGC::BF(weight_path, bias, N, L).infer(input, golden, num_inputs)   src/SDGC/BF/BF.hpp:55-68,102-124
 *   SNICIT_SDGC::SNIG(...).infer(...)                                            src/SDGC/SNIG/SNIG.hpp:51-66,101-131
 *   SNICIT_SDGC::XY(...).infer(...)                                              src/SDGC/XY/XY.hpp:41-54,85-201
 *   SNICIT_BEY::{SNICIT,BF,SNIG}(...).infer(input_path)                          src/beyond/SNICIT/SNICIT.hpp:59-74,134-172
 * Each entry point below names the part of those classes it replaces.  Ownership
 * follows the reference: a net owns its resident device weights (the reference
 * re-uploads them every round, SNICIT.hpp:252-317; here they stay in HBM), a plan
 * owns the activation buffers of one batch shape.  A context and everything made
 * from it must be used from one host thread at a time.
 */
#ifndef SNICIT_H
#define SNICIT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------ */
#define SNICIT_OK 0
#define SNICIT_E_INVALID (-1)   /* bad argument (sizes, mode, threshold ...)            */
#define SNICIT_E_HIP (-2)       /* HIP runtime error (message has the HIP error string)  */
#define SNICIT_E_NOMEM (-3)     /* device or host allocation failed / over budget       */
#define SNICIT_E_STATE (-4)     /* call out of order (e.g. run before set_input)        */
#define SNICIT_E_UNSUPPORTED (-5)
#define SNICIT_E_TIMEOUT (-6)       /* a wait ended before the request completed             */

/* ---- run modes (main/SDGC.cu:15-20,69-110; main/beyond.cu:12-18,104-150) --------- */
#define SNICIT_MODE_SNICIT 0      /* pre-threshold SpMM, clustering, delta SpMM, recovery */
#define SNICIT_MODE_BF 1          /* plain SpMM, all-zero rows skipped every layer         */
#define SNICIT_MODE_SNIG 2        /* plain SpMM, same results as BF                        */
#define SNICIT_MODE_XY 3          /* plain SpMM, same results as BF                        */
#define SNICIT_MODE_PLAIN_DENSE 4 /* plain SpMM over every row, no skipping                */

/* ---- bias forms -------------------------------------------------------------- */
#define SNICIT_BIAS_SCALAR 0      /* one value for the whole net (main/SDGC.cu:49-54)      */
#define SNICIT_BIAS_PER_LAYER 1   /* one value per layer: bias[L]                          */
#define SNICIT_BIAS_PER_NEURON 2  /* bias[L*N] (src/beyond/general.hpp:50,70)              */

/* ---- clustering variants ----------------------------------------------------- */
#define SNICIT_CLUSTER_EXACT 0    /* SDGC SNICIT: src/SDGC/SNICIT/kernel.hpp:8-218         */
#define SNICIT_CLUSTER_LOSSY 1    /* beyond SNICIT: src/beyond/SNICIT/kernel.hpp:5-183     */

typedef struct snicit_ctx snicit_ctx;
typedef struct snicit_net snicit_net;
typedef struct snicit_plan snicit_plan;

/* Run configuration.  snicit_run_cfg_default() fills the reference's values:
 * sample_size 32 and segments 16 (SNICIT.hpp:228-229), eps 0.03 (kernel.hpp:178,184),
 * lossy: pool 128 (main/beyond.cu:38), assign/delta tolerance 0.04 and post prune
 * 0.05 (src/beyond/SNICIT/kernel.hpp:71,90,161).  A tolerance e counts |x| > e with e
 * taken as the shortest decimal that rounds to the given float, as the reference compares
 * float magnitudes with its double literals (0.05f: |x| = 0.05f is above 0.05). */
typedef struct snicit_run_cfg {
  int32_t mode;            /* SNICIT_MODE_*                                            */
  int32_t threshold;       /* layer t at which clustering runs, 0 <= t < L (SNICIT)   */
  int32_t cluster;         /* SNICIT_CLUSTER_*                                         */
  int32_t sample_size;     /* exact: pool = first min(sample_size, live) live rows    */
  int32_t segments;        /* exact: segment sums per pool row (N % segments == 0)    */
  float eps;               /* exact: pruning tolerance on segment sums                */
  int32_t lossy_pool;      /* lossy: pool = lossy_pool strided rows                   */
  float lossy_eps_pool;    /* lossy: pruning tolerance between pool rows (0.03)       */
  float lossy_eps_assign;  /* lossy: L0 tolerance and delta pruning at clustering (0.04)*/
  float lossy_eps_post;    /* lossy: delta pruning after each post layer (0.05)       */
  int32_t use_graph;       /* 1: replay the layer loop as captured hipGraphs (plans   */
                           /* on nets with streamed weights always run directly)      */
  int32_t lossy_positive_only; /* lossy: a delta row is non-empty iff some entry > 0 at
                              * clustering (the reference's test, beyond/SNICIT/kernel.hpp:95);
                              * 0: iff some entry != 0                                    */
  int32_t reserved[6];
} snicit_run_cfg;

/* Per-run statistics.  Times are HIP-event milliseconds on the plan's stream with
 * inputs and weights resident, split like the reference's stage report
 * (SNICIT.hpp:659-664): pre = layers < t, cluster = clustering at t, post = layers
 * >= t, recover = recovery + categories.  Plain modes report everything in pre. */
typedef struct snicit_stats {
  double ms_total, ms_pre, ms_cluster, ms_post, ms_recover;
  double edges;              /* samples x nnz(W_l) x layers (nominal work)           */
  int64_t batch;             /* samples in the run                                    */
  int64_t live_at_threshold; /* B' (live rows entering clustering)                    */
  int64_t nerow_first;       /* non-empty delta rows after clustering                 */
  int64_t nerow_last;        /* non-empty delta rows after the last layer             */
  int64_t active;            /* samples with category 1                               */
  int32_t centroids;         /* k                                                     */
  int32_t layers;
  int64_t rows_processed;    /* sum over layers of rows the SpMM read                  */
  int64_t reserved[6];
} snicit_stats;

/* ---- context --------------------------------------------------------------------
 * Replaces the reference's GpuEnv (src/SDGC/gpu_lib/gpu_env.h:11-89): one device and one
 * HIP stream (default priority). */
int snicit_ctx_create(int device, snicit_ctx** out);
int snicit_ctx_destroy(snicit_ctx* ctx);

/* Panel grouping of rows with fewer than 32 inputs in the nets this context creates next
 * (the role of the reference's offline inspector, src/SDGC/inspector/gpu_block_scheduler.h:
 * 58-171, which merges row blocks by input reuse): NONE keeps one input set per panel;
 * CONSECUTIVE lets each group of rows join the panel before it when their input sets
 * together hold at most 32 inputs; SEARCH (default) seeds a panel and adds the groups that
 * share the most inputs with it, at any distance, then fills what room is left in order.
 * Every grouping gives the same results, bit for bit. */
#define SNICIT_GROUP_NONE 0
#define SNICIT_GROUP_CONSECUTIVE 1
#define SNICIT_GROUP_SEARCH 2
int snicit_ctx_set_grouping(snicit_ctx* ctx, int32_t grouping);
const char* snicit_last_error(void);
void snicit_run_cfg_default(snicit_run_cfg* cfg);

/* ---- nets -----------------------------------------------------------------------
 * snicit_net_create_ell replaces the weight ingestion of Base::_load_weight
 * (src/SDGC/base.hpp:89-141) and of the SNICIT/XY preprocessing (SNICIT.hpp:127-183):
 * row o of layer l has K nonzeros cols[(l*N+o)*K + r] (input neuron) with value
 * vals[(l*N+o)*K + r], i.e. Y_{l+1}[s,o] = f(sum_r vals*Y_l[s,cols] + b).  This is the
 * transpose of the reference's TSV line "i j v" (input i -> output j, utility.hpp:109-113).
 * Rows with fewer than K nonzeros pad with value 0.  The library packs rows that share
 * one input set into panels (the 16-outputs-share-32-inputs structure the reference's XY
 * kernels exploit, src/SDGC/XY/kernel.hpp:111-149) and keeps everything in HBM.
 * f(z) = min(max(z,0),clip) (src/SDGC/XY/kernel.hpp:5-7; clip 32 for SDGC, 1 for beyond). */
int snicit_net_create_ell(snicit_ctx* ctx, int32_t N, int32_t L, int32_t K,
                          const int32_t* cols, const float* vals, int32_t bias_kind,
                          const float* bias, float clip, snicit_net** out);
/* Hidden layers as CSR: row_ptr holds L*(N+1) absolute offsets into cols/vals (layer l's
 * rows at row_ptr[l*(N+1) ..]).  rows_are_inputs = 1 reads row i as input neuron i feeding
 * the output neurons in cols, the orientation of the reference's beyond files
 * (src/beyond/SNICIT/SNICIT.hpp:261-283, general.hpp:36-72); 0 reads rows as output neurons.
 * Rows are padded to the longest (value-0 entries) and stored as panels or wide ELL. */
int snicit_net_create_csr(snicit_ctx* ctx, int32_t N, int32_t L, const int64_t* row_ptr,
                          const int32_t* cols, const float* vals, int32_t rows_are_inputs,
                          int32_t bias_kind, const float* bias, float clip, snicit_net** out);

/* Seeded SDGC-like net (32 nnz/row of `weight`, banded+permuted 16-chunk pattern),
 * built directly in panel form with the weights filled on the device; no reference
 * counterpart (the reference reads the Graph Challenge TSVs, src/SDGC/utility.hpp:65-157). */
int snicit_net_generate_sdgc(snicit_ctx* ctx, int32_t N, int32_t L, uint64_t seed,
                             float weight, float bias, float clip, snicit_net** out);
/* Dense input layer (Nin -> N: y = min(max(W_in x + b_in, 0), in_clip), w_in[N][Nin]) and
 * dense output layer (N -> Nout: logits = W_out y + b_out, w_out[Nout][N]) around the hidden
 * layers, with category = argmax of the logits (first maximum).  Replaces the beyond nets'
 * dense_input / dense_output / check_acc (src/beyond/general.hpp:5-33,74-128).  Either size
 * may be 0.  Call before creating plans on the net. */
int snicit_net_set_io(snicit_net* net, int32_t Nin, const float* w_in, const float* b_in,
                      float in_clip, int32_t Nout, const float* w_out, const float* b_out);
/* Keep the hidden layers' weights in pinned host memory instead of device memory: each
 * layer's arrays move into one pinned block, and every plan on the net copies the blocks
 * into `buffers` device slots (2..8) as its layers run, the copy of the next layers
 * overlapping the current layer's kernels.  Replaces the reference's streamed weights:
 * BF's two device buffers refilled per layer from pinned memory (src/SDGC/BF/BF.hpp:
 * 168-178) and SNIG's weight-buffer copies inside its graph (src/SDGC/SNIG/SNIG.hpp:
 * 207-255).  Results are unchanged.  Dense input/output layers and per-neuron biases stay
 * in device memory.  Plans on the net issue their stages directly rather than as captured
 * graphs, so the copies run on a DMA engine beside the layer kernels; a batch still
 * captures one graph, in which the copies run in line with the kernels.  Call before
 * creating plans on the net; it cannot be undone. */
int snicit_net_stream_weights(snicit_net* net, int32_t buffers);
int snicit_net_destroy(snicit_net* net);
int snicit_net_info(const snicit_net* net, int32_t* N, int32_t* L, int32_t* K,
                    int64_t* panels_total, int64_t* device_bytes);

/* ---- plans (one batch shape + mode on one net) ----------------------------------
 * A plan holds the device activations of up to max_batch samples and the captured
 * layer-loop graphs.  It replaces the per-round buffers of SNICIT::_infer
 * (SNICIT.hpp:205-317) and BF::_infer (BF.hpp:157-220). */
int snicit_plan_create(snicit_net* net, const snicit_run_cfg* cfg, int64_t max_batch,
                       snicit_plan** out);
int snicit_plan_destroy(snicit_plan* plan);
/* Upload y0 (B rows of N floats, sample-major like the reference's input[b][n],
 * utility.hpp:268-293) into the plan's resident input buffer.  Values must be finite:
 * SNICIT_E_INVALID when one is inf or NaN (padded and merged panel slots multiply the
 * inputs a row does not use by 0, which is exact only for finite inputs). */
int snicit_plan_set_input(snicit_plan* plan, const float* y0, int64_t B);
/* Generate B seeded binary blob images (samples first_sample..first_sample+B-1) on the
 * device into the resident input buffer; equals snicit_gen_images on the host. */
int snicit_plan_set_input_images(snicit_plan* plan, uint64_t seed, int64_t first_sample,
                                 int64_t B);
/* Run the whole layer loop on the resident input.  Replaces SNICIT::_infer / BF::_infer. */
int snicit_plan_run(snicit_plan* plan, snicit_stats* stats);
/* Plain modes only: run layers [0, stop_layer) and stop (for layer-local checks). */
int snicit_plan_run_layers(snicit_plan* plan, int32_t stop_layer, snicit_stats* stats);
/* Categories of the last run: cat[b] = 1 iff any final activation of sample b > 0
 * (identify, src/SDGC/utility.hpp:194-212; recover, src/SDGC/SNICIT/kernel.hpp:197-218). */
int snicit_plan_categories(snicit_plan* plan, int32_t* cat);
/* Final activations of the last run as B x N floats (sample-major).  SNICIT mode returns
 * the recovered Yhat_i + Yhat_m(i) (src/SDGC/SNICIT/kernel.hpp:210). */
int snicit_plan_activations(snicit_plan* plan, float* y);
/* Introspection of the last run.  what: 0 live rows after each layer (int64[L]),
 * 1 non-empty delta rows after each post layer (int64[L], 0 before t), 2 pool rows
 * as original sample ids (int32[pool]), 3 pool segment sums (float[pool*segments]),
 * 4 centroids as original sample ids (int32[k]), 5 per-sample assignment (int32[B]:
 * centroid index, or -1 for rows dead before t), 6 output-layer logits (float[B*Nout]).
 * Returns the element count. */
int64_t snicit_plan_trace(snicit_plan* plan, int32_t what, void* out, int64_t cap);

/* Per-launch timing: replays the plan's layer loop eagerly with a hipEvent pair around
 * every SpMM launch on the plan's stream (the reference's per-layer event timing,
 * SNICIT.hpp:351-354).  For layer l < cap: ms[l] = duration of the layer's plain or
 * delta-row SpMM launch, rows[l] = columns it produced, cent_ms[l] = duration of the
 * centroid-row launch (post-threshold layers, else 0).  Any pointer may be null. */
int snicit_plan_profile(snicit_plan* plan, double* ms, int64_t* rows, double* cent_ms,
                        int64_t cap);

/* One-call form mirroring SNICIT::infer(input, golden, num_inputs, batch): rounds of
 * cfg batch over B samples, categories for all B (cat[b] = 0 or 1, int32).  Differences from
 * the reference: the golden check is the caller's (compare cat with the golden set, as
 * snicit_sdgc does), and there is no final-activation output; a plan returns those
 * (snicit_plan_activations) for the batch it holds. */
int snicit_infer(snicit_net* net, const snicit_run_cfg* cfg, const float* y0, int64_t B,
                 int64_t batch, int32_t* cat, snicit_stats* stats);

/* snicit_infer without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops
 * and harnesses that forbid hipMalloc).  Every device buffer of the run (activations,
 * lists, counters, staging) is carved from one static device array of 64 MiB per GPU, held
 * until the process ends; runs that need more return SNICIT_E_NOMEM before any kernel
 * launches (64 MiB holds, e.g., N = 1024 with batch <= ~3500, or N = 4096 with batch <= ~900).
 * It calls no hipMalloc or hipFree, captures no graphs and records no events (stage times
 * in stats are 0); nets with streamed weights are refused.  The array is shared by every
 * call on that GPU and keeps the last call's contents; the library clears what must start
 * at zero.  Calls from threads of this process run one at a time per GPU; calls that
 * overlap on one GPU in any other way race. */
int snicit_infer_noalloc(snicit_net* net, const snicit_run_cfg* cfg, const float* y0, int64_t B,
                         int64_t batch, int32_t* cat, snicit_stats* stats);

/* ---- autotuning of the layer-kernel launch shape ----------------------------------
 * The reference derives its slab width from the device's shared memory per block
 * (utility.hpp:17-41) and its XY layout from an offline inspector (src/SDGC/inspector/).
 * Here the launch shape of the layer kernels is measured on the plan's own net and input:
 * grid_target = workgroups a layer launch aims for (with the panel count it sets the tile of
 * one workgroup: one panel's rows by a run of 64-column tiles), narrow = column count up to
 * which the thread-per-(row, column) path runs, ell_lds_kb = LDS budget (KB, <= 160) of the input tile
 * that wide-row layers stage (0: gather inputs directly), ell_chunked = 1 to stage 64-column
 * tiles in chunks of input rows with a thread per row (layers of >= 128 rows), 2 (default)
 * for a wavefront per row over double-buffered chunks of 128 input rows (sparse hidden
 * layers).  Results are cached in a JSON file keyed by kernel kind, N, row length, rows per
 * panel, batch size (next power of two), input density (nonzero fraction of the plan's input
 * in halving steps), the net's panel grouping and whether the sample size was tuned. */
typedef struct snicit_tune_info {
  int64_t grid_target;
  int32_t narrow;
  int32_t ell_lds_kb;   /* wide-row layers: LDS input-tile budget, 0 = direct gathers */
  int32_t ell_chunked;  /* wide-row layers: 1 thread per row, 2 wavefront per row (chunked) */
  int32_t from_cache;   /* 1: taken from the cache file, nothing measured */
  double ms_default;    /* one layer launch with the default shape (0 when cached) */
  double ms_tuned;      /* one layer launch with the chosen shape */
  char key[96];
  /* post-threshold phase (SNICIT modes): workgroups per post layer launch, the delta-row
   * count up to which it takes one thread per (row, column) pair, and the post stage of a
   * whole run with the default and the chosen shape (0 when not tuned) */
  int64_t post_grid_target;
  int32_t post_narrow;
  double ms_post_default;
  double ms_post_tuned;
  /* whole runs (plans with their own timing events; 0 otherwise): the pre-threshold stage
   * (plain modes: the first min(L, 64) layers) and the whole run with the default settings
   * and with the chosen ones */
  double ms_pre_default, ms_pre_tuned;
  double ms_run_default, ms_run_tuned;
  int32_t post_form;    /* chosen post-layer form (snicit_plan_set_post_form) */
  int32_t sample_size;  /* run configuration's sample_size after tuning */
  int32_t placements;   /* placements of the activation buffers timed (SNICIT_TUNE_PLACEMENT) */
  int32_t placement;    /* the one kept: 0 = the plan's own buffers */
  int32_t compact_slack; /* chosen compaction slack (snicit_plan_set_compact_slack) */
  int32_t plain_tiles;   /* chosen tiles per wavefront (snicit_plan_set_plain_tiles) */
} snicit_tune_info;

/* Needs the plan's input (set_input); leaves the plan without a completed run.  cache_path
 * may be null (measure, do not store).  Tunes the pre-threshold launch shape on layer 0 at the
 * full batch, then on whole runs at the plan's own live counts: the pre-threshold grid, and
 * (SNICIT modes) the post-threshold grid, narrow-path limit and post-layer form at the plan's
 * own delta-row counts.  Every shape gives the same results, bit for bit.  Same as
 * snicit_plan_autotune_ex with flags 0. */
int snicit_plan_autotune(snicit_plan* plan, const char* cache_path, snicit_tune_info* info);
/* Options of snicit_plan_autotune_ex.
 * SNICIT_TUNE_SAMPLE (exact SNICIT clustering only): also choose the clustering sample size
 *   (the pool of Alg. 1, SNICIT.hpp:228) among 8, 16, 24, 32 and the configured one, by
 *   whole-run time.  A size is kept only when the run's categories equal those of the
 *   configured size on the input it was tuned on (a cached choice is reused under the same
 *   key without running the check again); the centroids, and with them the last bits of the
 *   recovered activations, may differ.
 * SNICIT_TUNE_PLACEMENT: also time the whole run (issued directly, without graphs) on up to
 *   four fresh allocations of the activation ping-pong buffers (while device memory allows)
 *   and keep the fastest.  Layer
 *   kernels stream these buffers at rates that depend on where the allocation lands in
 *   physical memory (C3: whole runs from 178 to 204 ms on different allocations of the same
 *   plan); results are unchanged.  Not cached: it is measured on every call. */
#define SNICIT_TUNE_SAMPLE 1u
#define SNICIT_TUNE_PLACEMENT 2u
int snicit_plan_autotune_ex(snicit_plan* plan, const char* cache_path, uint32_t flags,
                            snicit_tune_info* info);
int snicit_plan_set_shape(snicit_plan* plan, int64_t grid_target, int32_t narrow,
                          int32_t ell_lds_kb, int32_t ell_chunked);
/* Post-threshold layers: workgroups a launch aims for (default 2048) and the delta-row count
 * up to which a launch takes one thread per (row, column) pair (default 64). */
int snicit_plan_set_post_shape(snicit_plan* plan, int64_t post_grid_target,
                               int32_t post_narrow);
/* Post-threshold layers whose panels hold at most two distinct weight rows (every SDGC net):
 * 0 = the LDS-staged post kernel; 1 = the wavefront-local post kernel, each wavefront
 * computing its panel's centroid rows; 2 (default) = the wavefront-local kernel reading each
 * panel's centroid chain records, which the delta launch of the layer before computes in a
 * fifth wavefront per workgroup (compact post layout, resident weights; otherwise as 1).  All three give the same bits.
 * No counterpart in the reference (its post layer is one kernel, kernel.hpp:71-136). */
int snicit_plan_set_post_form(snicit_plan* plan, int32_t post_form);
/* Compaction of the live list after each plain layer (and SNICIT's layers before t): the
 * list is kept as it is when at most slack/1024 of its columns died in the layer (0..1023;
 * default 16), so the next layer reads its columns in place, aligned, and carries the dead
 * ones as zeros until a later compaction drops them.  The lists entering clustering and
 * after the last layer are always exact; results and reported live counts do not change.
 * No counterpart in the reference (BF compacts its row list every layer, BF.hpp:183-200). */
int snicit_plan_set_compact_slack(snicit_plan* plan, int32_t slack);
/* Plain layers on the vector kernel (panels whose rows repeat one weight row, every SDGC
 * net): 64-column tiles each wavefront gathers per iteration, 1, 2 (default) or 4.  Tiles
 * of one wavefront fetch a 128-byte line they share once; more tiles take more registers.
 * Results do not change.  No counterpart in the reference. */
int snicit_plan_set_plain_tiles(snicit_plan* plan, int32_t tiles);

/* ---- batches: many uneven batches over several nets in one graph launch ------------
 * The reference runs one net over equal rounds of `batch` samples, one round after another
 * (SNICIT.hpp:87-203, main/SDGC.cu:69-110).  A snicit_batch holds any number of batch
 * segments of different sizes over several resident nets.  Segments of one net share that
 * net's layer launches (their columns are concatenated, and the live rows and the non-empty
 * delta rows are compacted across segments), while the threshold-layer clustering runs per
 * segment: pool, centroids and nearest-centroid assignment only see the segment's own rows,
 * so every segment's results equal a run of that batch alone.  Each net's layer loop runs on
 * its own stream; the whole batch is captured once as one hipGraph and replayed per run. */
typedef struct snicit_batch snicit_batch;
typedef struct snicit_segment {
  int32_t net;     /* index into the nets array of snicit_batch_create */
  int32_t pad;
  int64_t samples; /* samples in this segment (> 0) */
} snicit_segment;

/* cfgs[i] is the run configuration of nets[i] (mode, threshold, clustering). */
int snicit_batch_create(snicit_ctx* ctx, snicit_net* const* nets, const snicit_run_cfg* cfgs,
                        int32_t n_nets, const snicit_segment* segs, int32_t n_segs,
                        snicit_batch** out);
int snicit_batch_destroy(snicit_batch* batch);
/* Host input of one segment: samples x (Nin or N) floats, sample-major. */
int snicit_batch_set_input(snicit_batch* batch, int32_t seg, const float* y0);
/* Device-rendered blob images of samples first_sample.. for one segment. */
int snicit_batch_set_input_images(snicit_batch* batch, int32_t seg, uint64_t seed,
                                  int64_t first_sample);
/* Runs every segment (all inputs must be set).  ms = wall time of the graph launch on the
 * device; per_net (nullable, n_nets entries) gets each net's counters (times zero). */
int snicit_batch_run(snicit_batch* batch, double* ms, snicit_stats* per_net);
int snicit_batch_categories(snicit_batch* batch, int32_t seg, int32_t* cat);
/* Final hidden activations of one segment, samples x N. */
int snicit_batch_activations(snicit_batch* batch, int32_t seg, float* y);
/* Autotunes every net's plan on the batch's inputs (all set); info: n_nets entries or null. */
int snicit_batch_autotune(snicit_batch* batch, const char* cache_path, snicit_tune_info* info);

/* ---- serving: continuous batching of requests over resident nets -------------------
 * The reference has no serving path; its driver runs fixed rounds (SNICIT.hpp:87-203).
 * A server keeps its nets resident and owns a worker thread and a set of slots per net.  A
 * slot holds device buffers for slot_samples samples in up to slot_segments requests, its
 * own stream, and captured stage graphs whose batch-segment table is device data, so any mix
 * of requests replays the same graphs.  Requests queue per net; the worker fills a free slot
 * from the queue head (FIFO) and launches it at once when no other slot of that net is busy,
 * when the group is full, or when its oldest request has waited group_window_ms; requests
 * that arrive while other groups run are admitted into the next free slot, so groups of one
 * net and of different nets overlap on the device.  Slots are allocated at creation up to
 * hbm_budget_bytes (at least one per net).  Clustering runs per request, except that with an
 * exact-clustering net consecutive requests smaller than cluster_min samples share one
 * clustering segment (the categories of exact SNICIT do not depend on the centroids). */
typedef struct snicit_server snicit_server;
typedef struct snicit_server_cfg {
  int64_t hbm_budget_bytes;  /* device memory for all slots                              */
  int64_t slot_samples;      /* samples per launch group (>= the largest request)         */
  int32_t slot_segments;     /* requests (clustering segments) per launch group           */
  int32_t max_slots_per_net; /* groups of one net in flight at once                       */
  int32_t cluster_min;       /* exact nets: smaller requests share a clustering segment   */
  int32_t warm;              /* 1: capture every slot's graphs at creation                */
  double group_window_ms;    /* longest wait for a group to fill while the net is busy    */
  int64_t group_samples;     /* a group closes at this many samples (0: slot_samples); a  */
                             /* larger single request still runs, alone                   */
  int64_t pinned_bytes;      /* pinned host memory that holds request inputs from submit  */
                             /* until their group is done (0: none); inputs that do not   */
                             /* fit are kept in pageable memory and staged at launch      */
  double round_ms;           /* > 0: fixed periods instead of the launch rule above: at   */
                             /* each period boundary the requests queued by then become   */
                             /* due, and every due request of a net launches in its next  */
                             /* free slot (window and group-fill triggers are not used)   */
} snicit_server_cfg;

typedef struct snicit_request_info {
  double ms_queue;          /* submit -> launch of its group                           */
  double ms_run;            /* launch -> results on the host                           */
  double ms_total;          /* submit -> results on the host                           */
  int64_t group_samples;    /* samples in the request's launch group                   */
  int32_t group_requests;   /* requests in that group                                  */
  int32_t slot;
} snicit_request_info;

typedef struct snicit_server_stats {
  int64_t requests_done, samples_done, groups;
  int64_t device_bytes;     /* allocated by the server's slots                          */
  int32_t slots;            /* slots over all nets                                      */
  int32_t queued;           /* requests waiting for a slot                              */
} snicit_server_stats;

void snicit_server_cfg_default(snicit_server_cfg* cfg);
int snicit_server_create(snicit_ctx* ctx, snicit_net* const* nets, const snicit_run_cfg* cfgs,
                         int32_t n_nets, const snicit_server_cfg* cfg, snicit_server** out);
/* Stops the worker after the groups in flight; pending requests are dropped. */
int snicit_server_destroy(snicit_server* server);
/* samples x (Nin or N) host floats, copied before return; values must be finite (not
 * checked here, unlike snicit_plan_set_input). */
int snicit_server_submit(snicit_server* server, int32_t net, const float* x, int64_t samples,
                         uint64_t* request_id);
/* Device-rendered blob images of samples first_sample.. (nets without an input layer). */
int snicit_server_submit_images(snicit_server* server, int32_t net, uint64_t seed,
                                int64_t first_sample, int64_t samples, uint64_t* request_id);
/* Waits up to timeout_ms (< 0: no limit) for the request; on success copies its categories
 * (samples int32) and releases it.  SNICIT_E_TIMEOUT when it is not done in time. */
int snicit_server_wait(snicit_server* server, uint64_t request_id, double timeout_ms,
                       int32_t* cat, snicit_request_info* info);
int snicit_server_stats_get(snicit_server* server, snicit_server_stats* stats);

/* ---- host-side seeded generators (splitmix64 streams) ---------------------------- */
/* cols[(l*N+o)*32 + r] of the SDGC-like net (ascending per row); all values = weight. */
int snicit_gen_sdgc_ell(int32_t N, int32_t L, uint64_t seed, int32_t* cols);
/* B x N binary blob images for samples first_sample.. (row-major pixels). */
int snicit_gen_images(int32_t N, uint64_t seed, int64_t first_sample, int64_t B, float* out);
/* Unstructured MLP: L hidden layers of N rows with K distinct random inputs each, weights
 * N(0, 2/K) (cols/vals [L][N][K], ascending per row); optional dense w_in [N][Nin] with
 * N(0, 2/Nin) and w_out [Nout][N] with N(0, 2/N). */
int snicit_gen_unstructured(int32_t Nin, int32_t N, int32_t L, int32_t K, int32_t Nout,
                            uint64_t seed, int32_t* cols, float* vals, float* w_in, float* w_out);
/* B x Nin features uniform in [0, 1) for samples first_sample.. */
int snicit_gen_features(int32_t Nin, uint64_t seed, int64_t first_sample, int64_t B, float* out);

#ifdef __cplusplus
}
#endif
#endif /* SNICIT_H */
