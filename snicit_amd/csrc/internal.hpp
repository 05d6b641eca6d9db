// Internal declarations shared by the host C++ and the HIP kernels of libsnicit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/snicit.h"

namespace snicit {

// Every failure inside the library throws Error; the C-ABI layer converts it to a
// status code and records the message for snicit_last_error().
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SNICIT_HIP(expr)                                                               \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::snicit::Error(SNICIT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + \
                                               " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define SNICIT_REQUIRE(cond, msg)                                                      \
  do {                                                                                 \
    if (!(cond)) throw ::snicit::Error(SNICIT_E_INVALID, std::string(msg));            \
  } while (0)

// Panel slots: every panel lists exactly kSlots input neurons (rows with fewer nonzeros
// are padded with value-0 slots).  32 = the fixed fan-in of the Graph Challenge nets.
constexpr int kSlots = 32;
// Maximum output rows per panel (one 32-row matrix-core tile); larger sharing groups are
// split.
constexpr int kMaxPanelRows = 32;
// Centroid capacity of the exact variant (pool <= 32 rows => k <= 32).
constexpr int kMaxCentroidsExact = 32;
// Centroid capacity of the lossy variant (pool of 128 strided rows).
constexpr int kMaxCentroidsLossy = 128;
// Input neurons per LDS chunk of the wavefront-per-row wide-row kernel (one chunk of a
// 64 x w column tile is kEllChunk * 256 * w bytes; two are resident, double-buffered).
constexpr int kEllChunk = 64;
// Entries per group of the wavefront-per-row kernel's stream (one 32-byte load).
constexpr int kEllGroup = 4;
inline int32_t ell_chunks(int32_t in_rows) { return (in_rows + kEllChunk - 1) / kEllChunk; }

// A fan-in panel: rows of one layer that read the same kSlots input neurons.
// Host form of one layer, used by the packer and the generator.
struct HostLayerPanels {
  std::vector<int32_t> pcols;  // [P][kSlots] input neuron per slot (ascending)
  std::vector<int32_t> pbeg;   // [P+1] first output row (in panel order) of each panel
  std::vector<int32_t> outs;   // [N] output neuron of each panel-order row
  std::vector<float> w;        // [N][kSlots] weight of each slot, panel order (may be empty
                               //  when the net is generated: all weights equal)
};

// Weight layouts of a layer on the device:
//  row-major  w[j][r]           (j = panel-order row), read by the vector-ALU kernel;
//  slot-major w[p][r][g], g<32  (row g of panel p, zero for g >= rows), the matrix-core
//             operand layout: one 32-float run per slot.

// Device arrays of one hidden layer, as the layer kernels read them: resident in the net's
// arrays, or in a plan's device slot holding the layer's streamed block.
struct LayerArrays {
  const int32_t* pcols = nullptr;  // panel layers: [P][kSlots]
  const int32_t* pbeg = nullptr;   //   [P + 1]
  const int32_t* outs = nullptr;   //   [N]
  const float* w = nullptr;        //   row-major or slot-major weights
  const uint32_t* psame = nullptr; //   [P] bit g: row g has the weights of row g - 1
  const int32_t* ecols = nullptr;  // wide-row layers: [N][K]
  const float* evals = nullptr;    //   [N][K]
  const int32_t* erng = nullptr;   //   [N][chunks + 1]
  const int2* eent = nullptr;      //   entry stream (+ one group of zero slack)
};

// Byte offsets of a layer's arrays inside its streamed block (-1: the layer has none).
struct BlockLayout {
  int64_t pcols = -1, pbeg = -1, outs = -1, w = -1, psame = -1, ecols = -1, evals = -1,
          erng = -1, eent = -1;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  int32_t grouping = SNICIT_GROUP_SEARCH;  // short-row panel grouping of nets created next
};

struct Net {
  Ctx* ctx = nullptr;
  int32_t N = 0, L = 0, K = 0;
  int32_t bias_kind = SNICIT_BIAS_SCALAR;
  std::vector<float> bias_host;  // scalar: [1]; per layer: [L]; per neuron: [L*N]
  float clip = 32.f;
  bool skip_dead = true;  // every bias <= 0 -> an all-zero row stays all-zero
  int32_t grouping = -1;  // panel grouping the rows were packed with (-1: generated net)
  // Panel data, all layers concatenated.
  std::vector<int64_t> panel_off;  // [L+1] first panel of each layer
  int32_t* d_pcols = nullptr;      // [Ptot][kSlots]
  int32_t* d_pbeg = nullptr;       // [Ptot + L] (P_l + 1 entries per layer)
  int32_t* d_outs = nullptr;       // [L][N]
  float* d_w = nullptr;            // per layer: row-major [N][kSlots] or slot-major [P][kSlots][32]
  // [Ptot] duplicate rows per panel: bit g set when panel row g has exactly (bit for bit) the
  // weights of row g - 1.  Rows of a panel share their inputs, so such a row's multiply-add
  // chain gives the bits of the row before it; the packer orders each panel's rows so that
  // rows with equal weights are adjacent.
  uint32_t* d_psame = nullptr;
  std::vector<uint8_t> dup_rows;   // [L] every panel of the layer repeats one weight row
  std::vector<uint8_t> few_rows;   // [L] every panel of the layer has at most two distinct rows
  std::vector<int64_t> w_off;      // [L] first float of each layer in d_w
  std::vector<uint8_t> mfma;       // [L] layer runs on the matrix cores (slot-major weights)
  // plain-epilogue launches of row-major panel layers run panel_plain_mfma_kernel
  bool plain_mfma = true;
  float* d_bias = nullptr;         // [L][N] when bias_kind == PER_NEURON
  // wide-ELL layers (rows with more than 32 nonzeros): rows padded to ell_K (multiple of 16)
  std::vector<int32_t> ell_K;      // [L] padded row length, 0 for panel layers
  std::vector<int64_t> ell_off;    // [L] first entry of the layer in d_ecols / d_evals
  int32_t* d_ecols = nullptr;
  float* d_evals = nullptr;
  // The same rows as a stream of (input neuron - first neuron of its chunk, weight bits)
  // pairs for the wavefront-per-row kernel: per row, per chunk of kEllChunk
  // input neurons, the row's entries of that chunk in ascending input order, padded to a
  // multiple of kEllGroup with value-0 entries that repeat the chunk's last real input (a
  // zero product leaves the chain unchanged).  Row o's entries of chunk c are
  // [rng[o][c], rng[o][c+1]) of the layer's stream; [N][nch+1] per wide layer.
  std::vector<int64_t> erng_off;   // [L] first int of the layer in d_erng
  std::vector<int64_t> eent_off;   // [L] first pair of the layer in d_eent
  int32_t* d_erng = nullptr;
  int2* d_eent = nullptr;
  // optional dense input (Nin -> N, clip in_clip) and output (N -> Nout, linear) layers,
  // src/beyond/general.hpp:5-33,74-102; rows zero-padded to Kin / Kout (multiples of 16)
  int32_t Nin = 0, Kin = 0, Nout = 0, Kout = 0;
  float in_clip = 1.f;
  float* d_win = nullptr;   // [N][Kin]
  float* d_wint = nullptr;  // [Kin][win_ld], the same weights transposed and zero-padded to
                            //   whole 64-row tiles (matrix-core operand)
  int32_t win_ld = 0;
  float* d_woutt = nullptr; // [Kout][wout_ld], the output weights likewise (32-row tiles)
  int32_t wout_ld = 0;
  float* d_bin = nullptr;   // [N]
  float* d_wout = nullptr;  // [Nout][Kout]
  float* d_bout = nullptr;  // [Nout]
  // Streamed hidden layers (snicit_net_stream_weights, the reference's double-buffered
  // weight copies, src/SDGC/BF/BF.hpp:168-178): layer l's arrays sit in one pinned host
  // block of blk_bytes[l] bytes at h_blocks + blk_off[l], laid out as blk[l]; the device
  // copies above are freed, and every plan copies the blocks into stream_buffers device
  // slots of blk_max bytes as its layers run.
  int32_t stream_buffers = 0;  // 0: every layer resident in device memory
  char* h_blocks = nullptr;
  std::vector<int64_t> blk_off, blk_bytes;
  std::vector<BlockLayout> blk;
  int64_t blk_max = 0;
  int64_t nnz_hidden = 0;   // nonzeros per hidden layer as given (edges per sample per layer)
  int64_t device_bytes = 0;
  int32_t live_plans = 0;   // plans sized for this net's input/output layers

  bool has_io() const { return Nin > 0 || Nout > 0; }
  bool streamed() const { return stream_buffers > 0; }
  bool is_ell(int l) const { return ell_K[l] > 0; }

  int64_t panels(int l) const { return panel_off[l + 1] - panel_off[l]; }
  float bias_scalar(int l) const {
    if (bias_kind == SNICIT_BIAS_SCALAR) return bias_host[0];
    if (bias_kind == SNICIT_BIAS_PER_LAYER) return bias_host[l];
    return 0.f;
  }
  const float* bias_vec(int l) const {
    return bias_kind == SNICIT_BIAS_PER_NEURON ? d_bias + (size_t)l * N : nullptr;
  }
};

// Layer l's arrays: the net's resident arrays (slot null), or the layer's block copied to
// the device at `slot` (streamed nets).
LayerArrays layer_arrays(const Net& net, int l, const char* slot);

// ---- host-side generators (gen.cpp) -------------------------------------------
uint64_t sm_mix(uint64_t z);
struct Rng {
  uint64_t s;
  uint64_t next();
};
void sdgc_layer_structure(int32_t N, int32_t l, uint64_t seed, int32_t* shift, int32_t* bit,
                          std::vector<int32_t>& perm);
void sdgc_layer_panels(int32_t N, int32_t l, uint64_t seed, HostLayerPanels& out);
void sdgc_layer_ell(int32_t N, int32_t l, uint64_t seed, int32_t* cols /*N*32*/);

struct ImageProto {
  int32_t cx, cy, r, pad;
};
struct ImageGeom {
  int32_t W, H, nproto_blobs[10], first_blob[10];
  std::vector<ImageProto> blobs;
};
ImageGeom image_geometry(int32_t N, uint64_t seed);
// Per-sample parameters: class, dx, dy, keep mask (bit b = blob b kept).
struct SampleParams {
  int32_t cls, dx, dy, keep;
};
SampleParams image_sample_params(const ImageGeom& g, uint64_t seed, int64_t sample);
void image_render(const ImageGeom& g, const SampleParams& p, int32_t N, float* row);

void unstructured_layer(int32_t N, int32_t K, int32_t l, uint64_t seed, int32_t* cols,
                        float* vals);
void dense_layer(int32_t rows, int32_t fan_in, uint64_t seed, uint64_t tag, float* w);
void uniform_features(int32_t Nin, uint64_t seed, int64_t first, int64_t B, float* out);

// ---- packing (net.cpp) -----------------------------------------------------------
// grouping: SNICIT_GROUP_NONE / _CONSECUTIVE / _SEARCH for rows with fewer than 32 inputs.
void pack_layer_ell(int32_t N, int32_t K, const int32_t* cols, const float* vals,
                    HostLayerPanels& out, int32_t grouping);

}  // namespace snicit
