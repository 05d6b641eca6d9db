// Seeded synthetic workloads: SDGC-like nets and binary blob images.
//
// The reference runs on the Graph Challenge Radix-Net files and MNIST-derived images
// fetched from S3 (bin/get_SDGC_dataset.sh:48-59), which are not available offline.
// These generators produce workloads of the same shape: exactly 32 nonzeros per row
// in groups of 16 outputs that share their 32 inputs (the structure the reference's XY
// kernels rely on, src/SDGC/XY/kernel.hpp:111-149, and that its layer-1..log2(N/16)
// stride pattern shows, src/SDGC/SNICIT/SNICIT.hpp:329-342), and sqrt(N) x sqrt(N)
// binary images.  Every random draw comes from a splitmix64 stream so the host and the
// device produce identical data.
#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace snicit {

uint64_t sm_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t Rng::next() {
  s += 0x9E3779B97F4A7C15ULL;
  return sm_mix(s);
}

static int32_t ilog2(int64_t v) {
  int32_t r = 0;
  while ((int64_t(1) << (r + 1)) <= v) ++r;
  return r;
}

// Layer l of the banded+permuted net: M = N/16 chunks of 16 neurons.  Output group g
// (neurons 16g..16g+15) reads chunks perm[x] and perm[x ^ 2^bit] with x = (g+shift) mod M,
// so every neuron has in-degree 32 and the two groups g, g' with x' = x ^ 2^bit share
// the same 32 inputs.
void sdgc_layer_structure(int32_t N, int32_t l, uint64_t seed, int32_t* shift, int32_t* bit,
                          std::vector<int32_t>& perm) {
  const int32_t M = N / 16;
  const int32_t lg = ilog2(M);
  Rng r{seed ^ sm_mix((uint64_t)(l + 1) * 0xD1B54A32D192ED03ULL)};
  *shift = (int32_t)(r.next() % (uint64_t)M);
  *bit = l % lg;
  perm.resize(M);
  for (int32_t i = 0; i < M; ++i) perm[i] = i;
  for (int32_t i = M - 1; i > 0; --i) {
    int32_t j = (int32_t)(r.next() % (uint64_t)(i + 1));
    std::swap(perm[i], perm[j]);
  }
}

static void check_sdgc_n(int32_t N) {
  SNICIT_REQUIRE(N >= 32 && (N & (N - 1)) == 0,
                 "SDGC-like nets need N a power of two >= 32, got " + std::to_string(N));
}

void sdgc_layer_ell(int32_t N, int32_t l, uint64_t seed, int32_t* cols) {
  check_sdgc_n(N);
  int32_t shift, bit;
  std::vector<int32_t> perm;
  sdgc_layer_structure(N, l, seed, &shift, &bit, perm);
  const int32_t M = N / 16;
  for (int32_t g = 0; g < M; ++g) {
    int32_t x = (g + shift) % M;
    int32_t a = perm[x], b = perm[x ^ (1 << bit)];
    int32_t lo = std::min(a, b), hi = std::max(a, b);
    for (int32_t j = 0; j < 16; ++j) {
      int32_t* row = cols + (size_t)(16 * g + j) * kSlots;
      for (int32_t r = 0; r < 16; ++r) row[r] = 16 * lo + r;
      for (int32_t r = 0; r < 16; ++r) row[16 + r] = 16 * hi + r;
    }
  }
}

void sdgc_layer_panels(int32_t N, int32_t l, uint64_t seed, HostLayerPanels& out) {
  check_sdgc_n(N);
  int32_t shift, bit;
  std::vector<int32_t> perm;
  sdgc_layer_structure(N, l, seed, &shift, &bit, perm);
  const int32_t M = N / 16;
  const int32_t P = M / 2;
  out.pcols.resize((size_t)P * kSlots);
  out.pbeg.resize(P + 1);
  out.outs.resize(N);
  out.w.clear();
  int32_t p = 0;
  for (int32_t x = 0; x < M; ++x) {
    if (x & (1 << bit)) continue;
    int32_t x2 = x | (1 << bit);
    int32_t a = perm[x], b = perm[x2];
    int32_t lo = std::min(a, b), hi = std::max(a, b);
    int32_t* pc = out.pcols.data() + (size_t)p * kSlots;
    for (int32_t r = 0; r < 16; ++r) pc[r] = 16 * lo + r;
    for (int32_t r = 0; r < 16; ++r) pc[16 + r] = 16 * hi + r;
    int32_t ga = ((x - shift) % M + M) % M;
    int32_t gb = ((x2 - shift) % M + M) % M;
    int32_t g0 = std::min(ga, gb), g1 = std::max(ga, gb);
    out.pbeg[p] = 32 * p;
    for (int32_t j = 0; j < 16; ++j) out.outs[32 * p + j] = 16 * g0 + j;
    for (int32_t j = 0; j < 16; ++j) out.outs[32 * p + 16 + j] = 16 * g1 + j;
    ++p;
  }
  out.pbeg[P] = N;
}

// ---- binary blob images ------------------------------------------------------------
// W x H raster (W = ceil(sqrt(N)), H = N / W; pixel n = y*W + x, pixels past W*H are 0).
// Ten prototypes of 3..6 discs with radius 8%..16% of min(W,H); a sample picks a
// prototype, shifts it by dx,dy in [-2,2] and drops each disc with probability 1/10.
ImageGeom image_geometry(int32_t N, uint64_t seed) {
  SNICIT_REQUIRE(N >= 16, "image size too small");
  ImageGeom g;
  g.W = (int32_t)std::ceil(std::sqrt((double)N));
  while ((int64_t)g.W * g.W < N) ++g.W;
  while (g.W > 1 && (int64_t)(g.W - 1) * (g.W - 1) >= N) --g.W;
  g.H = N / g.W;
  const int32_t m = std::min(g.W, g.H);
  const int32_t rmin = std::max(1, m * 8 / 100);
  const int32_t rmax = std::max(rmin, m * 16 / 100);
  Rng r{seed ^ 0x7F4A7C159E3779B9ULL};
  for (int c = 0; c < 10; ++c) {
    int32_t nb = 3 + (int32_t)(r.next() % 4);
    g.first_blob[c] = (int32_t)g.blobs.size();
    g.nproto_blobs[c] = nb;
    for (int b = 0; b < nb; ++b) {
      ImageProto p;
      p.cx = (int32_t)(r.next() % (uint64_t)g.W);
      p.cy = (int32_t)(r.next() % (uint64_t)g.H);
      p.r = rmin + (int32_t)(r.next() % (uint64_t)(rmax - rmin + 1));
      p.pad = 0;
      g.blobs.push_back(p);
    }
  }
  return g;
}

SampleParams image_sample_params(const ImageGeom& g, uint64_t seed, int64_t sample) {
  Rng r{seed ^ sm_mix((uint64_t)sample ^ 0x632BE59BD9B4E019ULL)};
  SampleParams p;
  p.cls = (int32_t)(r.next() % 10);
  p.dx = (int32_t)(r.next() % 5) - 2;
  p.dy = (int32_t)(r.next() % 5) - 2;
  p.keep = 0;
  for (int b = 0; b < g.nproto_blobs[p.cls]; ++b)
    if (r.next() % 10 != 0) p.keep |= 1 << b;
  return p;
}

void image_render(const ImageGeom& g, const SampleParams& p, int32_t N, float* row) {
  const int32_t nb = g.nproto_blobs[p.cls];
  const ImageProto* bl = g.blobs.data() + g.first_blob[p.cls];
  for (int32_t n = 0; n < N; ++n) {
    int32_t x = n % g.W, y = n / g.W;
    float v = 0.f;
    if (y < g.H) {
      for (int32_t b = 0; b < nb; ++b) {
        if (!((p.keep >> b) & 1)) continue;
        int32_t ddx = x - bl[b].cx - p.dx, ddy = y - bl[b].cy - p.dy;
        if (ddx * ddx + ddy * ddy <= bl[b].r * bl[b].r) {
          v = 1.f;
          break;
        }
      }
    }
    row[n] = v;
  }
}

// ---- unstructured nets (beyond-style MLPs) -----------------------------------------------
// Hidden layer l: row o draws K distinct inputs uniformly (partial Fisher-Yates from
// Rng{seed ^ mix(l*N + o + 1)}) with weights N(0, 2/K) by Box-Muller; the dense input layer
// (N x Nin) uses N(0, 2/Nin) and the output layer (Nout x N) N(0, 2/N), each row from its own
// stream.  Features are U[0,1).
static double unit(Rng& r) { return (double)(r.next() >> 11) * (1.0 / 9007199254740992.0); }

static float gauss(Rng& r, double sigma) {
  const double u1 = unit(r), u2 = unit(r);
  const double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
  return (float)(z * sigma);
}

void unstructured_layer(int32_t N, int32_t K, int32_t l, uint64_t seed, int32_t* cols,
                        float* vals) {
  std::vector<int32_t> idx(N);
  const double sigma = std::sqrt(2.0 / K);
  for (int32_t o = 0; o < N; ++o) {
    Rng r{seed ^ sm_mix((uint64_t)l * (uint64_t)N + (uint64_t)o + 1)};
    for (int32_t i = 0; i < N; ++i) idx[i] = i;
    for (int32_t i = 0; i < K; ++i) {
      const int32_t j = i + (int32_t)(r.next() % (uint64_t)(N - i));
      std::swap(idx[i], idx[j]);
    }
    std::sort(idx.begin(), idx.begin() + K);
    for (int32_t i = 0; i < K; ++i) {
      cols[(size_t)o * K + i] = idx[i];
      vals[(size_t)o * K + i] = gauss(r, sigma);
    }
  }
}

void dense_layer(int32_t rows, int32_t fan_in, uint64_t seed, uint64_t tag, float* w) {
  const double sigma = std::sqrt(2.0 / fan_in);
  for (int32_t o = 0; o < rows; ++o) {
    Rng r{seed ^ sm_mix(tag ^ ((uint64_t)o + 1) * 0x9E3779B97F4A7C15ULL)};
    for (int32_t i = 0; i < fan_in; ++i) w[(size_t)o * fan_in + i] = gauss(r, sigma);
  }
}

void uniform_features(int32_t Nin, uint64_t seed, int64_t first, int64_t B, float* out) {
  for (int64_t b = 0; b < B; ++b) {
    Rng r{seed ^ sm_mix((uint64_t)(first + b) ^ 0x3C6EF372FE94F82AULL)};
    for (int32_t i = 0; i < Nin; ++i) out[(size_t)b * Nin + i] = (float)unit(r);
  }
}

}  // namespace snicit
