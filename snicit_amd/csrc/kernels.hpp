// Launch wrappers of the gfx950 kernels (kernels.hip).  Activations live feature-major:
// Y[n * ld + pos] holds neuron n of the sample at column pos, so 64 consecutive samples
// (one wavefront) read or write one 256-byte run of a neuron row.
#pragma once

#include "internal.hpp"

namespace snicit {

enum Epilogue : int {
  EPI_PLAIN = 0,        // y = f(z + b); flag = any(y > 0)
  EPI_CENTROID = 1,     // zc = z + b stored; y = f(zc)
  EPI_DELTA = 2,        // y = f(zc[m] + z) - f(zc[m]); flag = any(y != 0)
  EPI_DELTA_LOSSY = 3,  // as DELTA, then |y| <= eps -> 0
  EPI_LINEAR = 4,       // y = z + b (dense output layer, no activation)
};

// Batch segments of a plan's columns, each clustered on its own.  Centroid column
// s*stride + c is centroid c of segment s (c < kseg[s]); kcols = columns in use
// (s*stride + kseg[s] of the last segment with centroids).  Segments may be empty.
struct Segs {
  int32_t nseg;
  int32_t stride;
  int32_t* seglo;        // device [nseg+1]
  int32_t* kseg;         // device [nseg]
  int32_t* kcols;        // device [1]
};

// Tolerances.  The reference writes them as double literals (0.03, 0.04, 0.05) and
// compares float magnitudes with them in double (src/SDGC/SNICIT/kernel.hpp:178,184;
// src/beyond/SNICIT/kernel.hpp:31,38,90,161).  A tolerance arrives as a float; its value is
// the shortest decimal that rounds to that float (tol_decimal: 0.05f -> 0.05), and the
// kernels compare |x| > tol_cut(e), the largest float not above that decimal, which for
// every float x equals (double)|x| > tol_decimal(e).
double tol_decimal(float e);
float tol_cut(float e);

// Row stride of a post-threshold delta buffer that holds c list positions (compact post
// layout): neuron n of position j sits at n * compact_ld(c) + j, rows 128-byte aligned.
// Against the plan's batch-capacity stride this keeps a layer's delta rows in one dense
// region (65536 rows of 2 KB instead of rows 240 KB apart at N = 65536, 485 rows).
__host__ __device__ inline int64_t compact_ld(int64_t c) { return ((c > 0 ? c : 1) + 31) / 32 * 32; }
// Compact centroid stride of the post stage: k rounded up to a power of two (at most kcap).
__host__ __device__ inline int compact_kst(int k, int kcap) {
  int s = 1;
  while (s < k && s < kcap) s <<= 1;
  return s < kcap ? s : kcap;
}

// Ring of per-layer non-empty flags of the compact post layout: post layer j writes slot
// j % kPostRing, and a compaction at most every kPostRing layers counts and clears them.
constexpr int kPostRing = 64;

// Launch shape of the layer kernels (chosen per net, batch size and density by the autotuner):
// grid_target = workgroups a layer launch aims for (panels x column tiles), narrow = column
// count up to which the thread-per-(row, column) path runs.
struct LaunchShape {
  int64_t grid_target = 16384;
  int32_t narrow = 64;
  // workgroups a post-threshold layer launch aims for (panels x 256-column ranges); each
  // workgroup stages its panel's weights and centroid inputs once for all of its columns,
  // so fewer, longer workgroups read less
  int64_t post_grid_target = 2048;
  // delta-row count up to which the post kernel takes one thread per (row, column) pair
  int32_t post_narrow = 64;
  // post layers whose panels hold at most two distinct weight rows: 0 = the LDS-staged post
  // kernel; 1 = the wavefront-local post kernel (no LDS, no barrier), computing its panel's
  // centroid chains itself; 2 = the wavefront-local kernel reading chain records that the
  // launch of the layer before computed in a fifth wavefront (compact post layout, resident
  // weights; else 1)
  int32_t post_form = 2;
  // plain layers (and SNICIT's layers < t): the compaction after a layer keeps the live list
  // as it is when at most compact_slack / 1024 of it died, so the next layer reads aligned
  // columns (a shifted list costs ~12% more reads) and carries the dead columns as zeros;
  // the lists entering clustering and the last layer's are always exact
  int32_t compact_slack = 16;
  // plain layers on the vector kernel (panels whose rows repeat one weight row): 64-column
  // tiles per wavefront and iteration, 1, 2 or 4
  int32_t plain_tiles = 2;
  // LDS budget (KB) of the staged wide-row kernel's input tile; 0 = direct-gather kernel
  int32_t ell_lds_kb = 80;
  // wide-row layers with >= 128 rows: 1 = thread per row over 64-column tiles staged in
  // chunks of input rows; 2 = wavefront per row, lanes = 64 columns, input chunks of
  // kEllChunk rows double-buffered in LDS (layers with a chunk range table), and dense
  // layers with a transposed weight copy on the matrix cores; 0 = neither
  int32_t ell_chunked = 2;
};

struct SpmmLaunch {
  const Net* net;
  int layer;
  LayerArrays la;        // the layer's device arrays
  int epilogue;
  const float* yin;
  int64_t ldin;
  float* yout;
  int64_t ldout;
  const int32_t* count;  // device: number of columns to produce
  const int32_t* map;    // device: input column of each output column (null = identity)
  int32_t* flags;        // device: per output column flag (null = none)
  float* zc;             // centroid pre-activations [N][kcap]
  int32_t kcap;
  const int32_t* mcent;  // per output column centroid index (DELTA)
  float eps_post;
  int64_t max_cols;      // upper bound of *count (grid sizing)
  LaunchShape shape;
};
void launch_panel_spmm(const SpmmLaunch& a, hipStream_t s);

// One post-threshold layer in one launch: centroid rows (cin -> cout, k columns of ld kcap)
// and delta rows (din -> dout through map, centroid index mcent).  Needs row-major weights
// (post_layer_fused); otherwise the centroid and delta epilogues of launch_panel_spmm run.
// Compact post layout: ldin_cnt (nullable) gives din's stride compact_ld(*ldin_cnt),
// compact_out writes dout with stride compact_ld(*count), and nlist_next (nullable) receives
// *count for the next layer.
bool post_layer_fused(const Net& net, int l);
// The next post layer's chain records and centroid rows, computed inside a delta launch
// (pbeg null: none): that layer's panel arrays, bias, centroid columns in and out, records.
struct NextChain {
  const int32_t* pcols = nullptr;
  const int32_t* pbeg = nullptr;
  const int32_t* outs = nullptr;
  const float* w = nullptr;
  const uint32_t* psame = nullptr;
  const float* bias_vec = nullptr;
  float bias_s = 0.f;
  const float* cin = nullptr;
  float* cout = nullptr;
  float* zchain = nullptr;
  int64_t P = 0;
};
void launch_panel_post(const Net& net, int l, const LayerArrays& la, bool lossy,
                       const float* cin, float* cout, int32_t kcap, const int32_t* k,
                       const float* din, int64_t ldin, const int32_t* ldin_cnt, float* dout,
                       int64_t ldout, bool compact_out, const int32_t* count, int32_t* nlist_next,
                       const int32_t* map, const int32_t* mcent, int32_t* flags, float eps_post,
                       int64_t max_cols, const LaunchShape& shape, hipStream_t s,
                       bool ccompact = false, const float* zchain = nullptr,
                       const NextChain& nx = NextChain());
// Chain records of a post layer (floats per panel, see post_chain_kernel) and its centroid
// rows cin -> cout, apart from the delta rows; layers with net.few_rows[l] only.
constexpr int kPostChainRec = 64;
void launch_post_chain(const Net& net, int l, const LayerArrays& la, const float* cin,
                       float* cout, int32_t kcap, const int32_t* k, bool ccompact, float* zchain,
                       hipStream_t s);
// Centroid columns (k of them, N rows) copied between the kcap stride and the compact stride
// compact_kst(k, kcap) of the post stage: src_compact / dst_compact pick each side's stride.
void launch_centroid_restride(const float* src, bool src_compact, float* dst, bool dst_compact,
                              int32_t kcap, const int32_t* k, int64_t N, hipStream_t s);
// Centroid rows of a post-threshold layer: zc = W cin + b, cout = f(zc) (k columns, ld kcap).
void launch_panel_centroid(const Net& net, int l, const LayerArrays& la, const float* cin,
                           float* cout, float* zc, int32_t kcap, const int32_t* k,
                           hipStream_t s);

// Wide ELL layer (rows with more than 32 nonzeros, or a dense layer when cols is null):
// rows x K entries, ascending input per row, zero-padded; yin has any number of rows.
struct EllLaunch {
  const int32_t* cols;  // [rows][K] or null (column r = input r: dense layer)
  const float* vals;    // [rows][K]
  const int32_t* rng;   // [rows][ell_chunks(in_rows)+1] ranges in ent, or null
  const int2* ent;      // padded (chunk tile byte offset, weight bits) stream, with rng
  const float* wt;      // dense layers: [in_rows][wt_ld] transposed weights, or null
  int32_t wt_ld;        //   rows padded with zero weights to a multiple of 32
  int32_t K, rows;
  int32_t in_rows;      // input neurons the entries refer to (rows of yin read)
  const float* bias_vec;
  float bias_s, clip;
  int epilogue;
  const float* yin;
  int64_t ldin;
  float* yout;
  int64_t ldout;
  const int32_t* count;
  const int32_t* map;
  int32_t* flags;
  float* zc;
  int32_t kcap;
  const int32_t* mcent;
  float eps_post;
  int64_t max_cols;
  LaunchShape shape;
};
void launch_ell_spmm(const EllLaunch& a, hipStream_t s);
// cat[j] = argmax_o logits[o][j] (first maximum, starting from -10000), j < *count
void launch_argmax(const float* logits, int64_t ld, int32_t rows, const int32_t* count,
                   int64_t max_cols, int32_t* cat, hipStream_t s);

// ---- lossy clustering (src/beyond/SNICIT/kernel.hpp:5-101) ---------------------------
// pool row y of a segment of B_s columns = its column y*B_s/pool, gathered sample-major:
// poolrows [nseg*pool][N], poolcol [nseg*pool] (column, -1 in an empty segment)
void launch_lossy_pool(const float* Y, int64_t ld, const Segs& sg, int32_t pool, int32_t N,
                       float* poolrows, int32_t* poolcol, hipStream_t s);
// close: scratch of [nseg][128][2] words (each pool row's mask of close later rows)
void launch_lossy_select(const float* poolrows, const int32_t* poolcol, const Segs& sg,
                         int32_t pool, int32_t N, float eps, uint64_t* close, int32_t* cent_col,
                         hipStream_t s);
// dist: [128][ld] int32 scratch (centroid-major, any contents on entry); centmark[i] = 1
// for centroids
void launch_lossy_assign(const float* Y, int64_t ld, const Segs& sg, int64_t max_cols,
                         const float* C, int32_t kcap, const int32_t* cent_col, int32_t N,
                         float eps, int32_t* dist, int32_t* mcar, int32_t* centmark, int32_t* rec,
                         hipStream_t s);
// D[n*ldD + i] = |Y - C_m| > eps ? Y - C_m : 0 for non-centroid columns; flags = any(v != 0)
// (or any(v > 0) with positive_only, the reference's test)
void launch_lossy_delta(const float* Y, int64_t ld, const float* C, int32_t kcap,
                        const int32_t* mcar, const int32_t* centmark, const int32_t* B,
                        int64_t max_cols, int32_t N, float eps, bool positive_only, float* D,
                        int64_t ldD, int32_t* flags, hipStream_t s);

// max_cols: upper bound of *count_in; tile_cnt: scratch of compact_tiles(max_cols) ints, or
// null for lists expected to stay short (one workgroup then walks any count, one launch).
// slack (0..1023): keep the list as it is when at most count * slack / 1024 of its columns
// are unflagged (0: always compact); live_out (nullable) receives the flagged count.
void launch_compact(int32_t* flags, const int32_t* count_in, int32_t* count_out, int32_t* map_out,
                    const int32_t* compose, const int32_t* c0_in, int32_t* c0_out,
                    const int32_t* c1_in, int32_t* c1_out, int64_t max_cols, int32_t* tile_cnt,
                    hipStream_t s, int32_t slack = 0, int32_t* live_out = nullptr);
inline int64_t compact_tiles(int64_t max_cols) { return (max_cols + 4095) / 4096; }

// nerow[i] = set flags of ring slot (j0 + i) % kPostRing over *count positions, i < nslots;
// every slot but the last is cleared
void launch_ring_count(int32_t* ring, int64_t ring_ld, int32_t j0, int32_t nslots,
                       const int32_t* count, int32_t* nerow, hipStream_t s);
void launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
void launch_iota_i32(int32_t* p, int64_t n, hipStream_t s);
void launch_fill_f32(float* p, float v, int64_t n, hipStream_t s);
// adds the count of nonzero entries of y[0..rows) x [0..B) (row pitch ld) to *out
void launch_count_nonzero(const float* y, int64_t rows, int64_t ld, int64_t B,
                          unsigned long long* out, hipStream_t s);

// [B][N] row-major (device) -> Y[n*ld + b]
// nonfinite (nullable): set to 1 when any value read is inf or NaN.
void launch_transpose_in(const float* src, int64_t B, int32_t N, float* dst, int64_t ld,
                         int32_t* nonfinite, hipStream_t s);
// Y[n*ld + b] -> [B][N] row-major (device), b < B
void launch_transpose_out(const float* src, int64_t ld, int64_t B, int32_t N, float* dst,
                          hipStream_t s);
// dst[n*lddst + dstcol[j]] = src[n*ldsrc + (map ? map[j] : j)], j < *count
void launch_scatter_cols(const float* src, int64_t ldsrc, const int32_t* map,
                         const int32_t* dstcol, const int32_t* count, int64_t max_cols,
                         int32_t N, float* dst, int64_t lddst, hipStream_t s);

void launch_images(const ImageGeom& g, const int32_t* d_params /*B x 4*/,
                   const int32_t* d_blobs /*nblobs x 4*/, const int32_t* d_proto /*20*/,
                   int64_t B, int32_t N, float* Y, int64_t ld, hipStream_t s);

// flags[b] = any(Y[:, b] != 0), b < *count
void launch_row_any(const float* Y, int64_t ld, const int32_t* count, int64_t max_cols,
                    int32_t N, int32_t* flags, hipStream_t s);
// cat[orig[j]] = 1 for j < *count (cat pre-filled with 0)
void launch_cat_from_list(const int32_t* orig, const int32_t* count, int64_t max_cols,
                          int32_t* cat, hipStream_t s);
// cat[j] = flags[j] != 0, flags reset, j < *count
void launch_cat_from_flags(int32_t* flags, const int32_t* count, int64_t max_cols, int32_t* cat,
                           hipStream_t s);

// ---- exact clustering (src/SDGC/SNICIT/kernel.hpp:8-69,138-194) -------------------
// seglo[s] = first list position of batch segment s (lower bound of segstart[s] in the
// ascending orig list), seglo[nseg] = count.
void launch_segment_bounds(const int32_t* orig, const int32_t* count, const int32_t* segstart,
                           const Segs& sg, hipStream_t s);
// F: [nseg][pool][segments]
void launch_seg_sums(const float* Y, int64_t ld, const int32_t* map, const Segs& sg, int32_t pool,
                     int32_t segments, int32_t N, float* F, hipStream_t s);
void launch_select_exact(const float* F, const Segs& sg, int32_t pool, int32_t segments, float eps,
                         int32_t* cent_slot, hipStream_t s);
void launch_gather_centroids(const float* Y, int64_t ld, const int32_t* map,
                             const int32_t* cent_slot, const Segs& sg, int32_t N, float* C,
                             int32_t kcap, hipStream_t s);
// dist: [max_cols][32] int32, all zero on entry and left zero on exit.
void launch_assign_exact(const float* Y, int64_t ld, const int32_t* map, const Segs& sg,
                         int64_t max_cols, const float* C, int32_t kcap, const int32_t* cent_slot,
                         int32_t N, int32_t* dist, int32_t* flags, int32_t* mcar,
                         const int32_t* orig, int32_t* rec, hipStream_t s);
// D[n*ldD + j] = Y[n*ld + map[j]] - C[n*kcap + m[j]]; compact: ldD = compact_ld(*count)
void launch_make_delta(const float* Y, int64_t ld, const int32_t* map, const float* C,
                       int32_t kcap, const int32_t* m, const int32_t* count, int64_t max_cols,
                       int32_t N, float* D, int64_t ldD, bool compact, hipStream_t s);

// ---- recovery (src/SDGC/SNICIT/kernel.hpp:197-218) ---------------------------------
void launch_centroid_cats(const float* C, int32_t kcap, const int32_t* k, int32_t N,
                          int32_t* catc, hipStream_t s);
void launch_recover_base(const int32_t* rec, const int32_t* B, int64_t max_cols,
                         const int32_t* catc, int32_t* cat, hipStream_t s);
// ldD_cnt (nullable): D has the compact post layout, stride compact_ld(*ldD_cnt)
void launch_recover_rows(const float* D, int64_t ldD, const int32_t* ldD_cnt, const int32_t* map,
                         const int32_t* m, const int32_t* orig, const int32_t* count,
                         int64_t max_cols, const float* C, int32_t kcap, int32_t N, int32_t* act,
                         int32_t* cat, hipStream_t s);
// Yfull[n*ld + b] = rec[b] < 0 ? 0 : C[n*kcap + rec[b]]
void launch_expand_centroids(const int32_t* rec, const int32_t* B, int64_t max_cols,
                             const float* C, int32_t kcap, int32_t N, float* Y, int64_t ld,
                             hipStream_t s);
// Y[n*ld + orig[j]] = D[n*ldD + map[j]] + C[n*kcap + m[j]] (ldD_cnt as in recover_rows)
void launch_scatter_delta(const float* D, int64_t ldD, const int32_t* ldD_cnt, const int32_t* map,
                          const int32_t* m, const int32_t* orig, const int32_t* count,
                          int64_t max_cols, const float* C, int32_t kcap, int32_t N, float* Y,
                          int64_t ld, hipStream_t s);

}  // namespace snicit
