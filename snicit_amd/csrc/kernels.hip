// gfx950 kernels of the SNICIT layer loop.
//
// Layout: activations are feature-major, Y[n * ld + pos] (neuron n of the sample held in
// column pos).  A wavefront owns 64 consecutive columns, so each neuron access is one
// contiguous 256-byte run; weights, panel inputs and output ids are wavefront-uniform and
// are fetched with scalar loads.  The sparse layers are bound by HBM bandwidth (one read and
// one write of every live activation per layer).  Panels whose rows all repeat one weight row
// (every SDGC net) need one chain per column and run it on the vector ALU
// (panel_spmm_tiles_kernel); plain panels with distinct rows and the dense input and output
// layers run their fixed-order multiply-add chains on the matrix cores
// (v_mfma_f32_32x32x1_2b_f32: with K = 1 each element is one fused multiply-add, so the
// chain's bits are those of the vector form); everything else runs on the vector ALU.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "kernels.hpp"

namespace snicit {

namespace {

constexpr int kBlock = 256;   // 4 wavefronts of 64
constexpr int kWave = 64;

__device__ __forceinline__ float act_clip(float z, float clip) {
  // f(z) = min(max(z, 0), clip): src/SDGC/XY/kernel.hpp:5-7, src/SDGC/BF/kernel.hpp:79-82.
  // One v_med3_f32 (equal to fminf(fmaxf(z, 0), clip) for every non-NaN z when clip >= 0).
  return __builtin_amdgcn_fmed3f(z, 0.f, clip);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// 256-column tiles of a launch over `count` list positions, split into gridDim.y contiguous
// runs: workgroup row blockIdx.y takes positions [p0, p1).  Neighbouring tiles share a
// 128-byte line wherever the column list is shifted; in one run they are read back to back
// by one workgroup, so the line is fetched once.
__device__ __forceinline__ void tile_run(int64_t count, int64_t& p0, int64_t& p1) {
  const int64_t tiles = (count + kBlock - 1) / kBlock;
  const int64_t per = (tiles + gridDim.y - 1) / gridDim.y;
  const int64_t t0 = min(tiles, (int64_t)blockIdx.y * per);
  p0 = t0 * kBlock;
  p1 = min(count, (t0 + per) * kBlock);
}

// Value of v in lane l (l wavefront-uniform).
__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ int lane_prefix(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// Column counts up to this use the narrow path (one thread per output x column).
// Dynamic LDS a workgroup may use (gfx950: 160 KB per workgroup, no attribute needed).
constexpr int64_t kLdsPerBlock = 160 * 1024;
constexpr int kNarrowCols = 64;  // default column count up to which narrow paths run

// Epilogue of one output value (see panel_spmm_kernel).
template <int EPI>
__device__ __forceinline__ float epilogue(float acc, float b, float clip, bool valid, int o,
                                          int64_t pos, int32_t mi, float* __restrict__ zc,
                                          int32_t kcap, float eps_post, bool& flag) {
  float y;
  if (EPI == EPI_PLAIN) {
    y = act_clip(acc + b, clip);
    flag |= (y > 0.f);
  } else if (EPI == EPI_CENTROID) {
    const float z = acc + b;
    if (valid) zc[(int64_t)o * kcap + pos] = z;
    y = act_clip(z, clip);
  } else if (EPI == EPI_LINEAR) {
    y = acc + b;
  } else {
    const float z = valid ? zc[(int64_t)o * kcap + mi] : 0.f;
    y = act_clip(z + acc, clip) - act_clip(z, clip);
    if (EPI == EPI_DELTA_LOSSY) y = fabsf(y) > eps_post ? y : 0.f;
    flag |= (y != 0.f);
  }
  return y;
}

// Accumulator of one v_mfma_f32_32x32x1_2b_f32 (two 32 x 32 blocks, 32 values per lane).
typedef float floatx32 __attribute__((ext_vector_type(32)));

// ---- panel SpMM ---------------------------------------------------------------------
// Block (panel p, column tile): each wavefront gathers the panel's 32 inputs of its 64
// columns into registers, then produces every output row of the panel with a fixed-order
// fused multiply-add chain over the slots (ascending input neuron):
//   acc = fma(w[31], x[31], ... fma(w[0], x[0], 0)).
// Plain (src/SDGC/BF/kernel.hpp:57-88): y = f(acc + b), flag = any(y > 0).
// Centroid / delta (src/SDGC/SNICIT/kernel.hpp:71-136): zc = acc + b for centroid rows;
// delta rows y = f(zc[m] + acc) - f(zc[m]), flag = any(y != 0).
// Lossy delta (src/beyond/SNICIT/kernel.hpp:141-168): |y| <= eps -> 0.
// A wavefront whose 64 x 32 inputs are all zero skips the multiply-adds (acc = +0 exactly,
// what the chain would give), the zero-input skipping of src/SDGC/SNICIT/kernel.hpp:96.
// With at most `narrow` columns the wavefront-per-64-columns layout would leave one
// wavefront walking the panel's rows behind dependent scalar loads, so those launches run
// one thread per (row, column) pair instead, with the same chain and epilogue.
// (row, column) pairs of one panel over every workgroup of the panel: each pair's 32 inputs
// are loaded together, then the chain runs.
template <int EPI>
__device__ __forceinline__ void panel_pairs(
    int beg, int end, const int32_t* __restrict__ pc, const int32_t* __restrict__ outs,
    const float* __restrict__ w, const float* __restrict__ bias_vec, float bias_s, float clip,
    const float* __restrict__ yin, int64_t ldin, float* __restrict__ yout, int64_t ldout,
    int64_t count, const int32_t* __restrict__ map, int32_t* __restrict__ flags,
    float* __restrict__ zc, int32_t kcap, const int32_t* __restrict__ mcent, float eps_post) {
  const int G = end - beg;
  const int64_t pairs = (int64_t)G * count;
  for (int64_t q = (int64_t)blockIdx.y * kBlock + threadIdx.x; q < pairs;
       q += (int64_t)gridDim.y * kBlock) {
    const int j = beg + (int)(q % G);
    const int64_t pos = q / G;
    const int64_t src = map ? (int64_t)map[pos] : pos;
    const float4* __restrict__ w4 = reinterpret_cast<const float4*>(w + (size_t)j * kSlots);
    float x[kSlots];
#pragma unroll
    for (int r = 0; r < kSlots; ++r) x[r] = yin[(int64_t)pc[r] * ldin + src];
    float acc = 0.f;
#pragma unroll
    for (int r4 = 0; r4 < kSlots / 4; ++r4) {
      const float4 wv = w4[r4];
      acc = __builtin_fmaf(wv.x, x[4 * r4 + 0], acc);
      acc = __builtin_fmaf(wv.y, x[4 * r4 + 1], acc);
      acc = __builtin_fmaf(wv.z, x[4 * r4 + 2], acc);
      acc = __builtin_fmaf(wv.w, x[4 * r4 + 3], acc);
    }
    const int o = outs[j];
    const float b = bias_vec ? bias_vec[o] : bias_s;
    const int32_t mi = (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) ? mcent[pos] : 0;
    bool flag = false;
    const float y = epilogue<EPI>(acc, b, clip, true, o, pos, mi, zc, kcap, eps_post, flag);
    yout[(int64_t)o * ldout + pos] = y;
    if (flags && flag) flags[pos] = 1;
  }
}

template <int EPI>
__global__ __launch_bounds__(kBlock, 8) void panel_spmm_kernel(
    const int32_t* __restrict__ pcols, const int32_t* __restrict__ pbeg,
    const int32_t* __restrict__ outs, const float* __restrict__ w,
    const float* __restrict__ bias_vec, float bias_s, float clip, const float* __restrict__ yin,
    int64_t ldin, float* __restrict__ yout, int64_t ldout, const int32_t* __restrict__ count_ptr,
    const int32_t* __restrict__ map, int32_t* __restrict__ flags, float* __restrict__ zc,
    int32_t kcap, const int32_t* __restrict__ mcent, float eps_post, int32_t narrow,
    const uint32_t* __restrict__ psame) {
  const int p = blockIdx.x;
  const int64_t count = *count_ptr;
  const int beg = pbeg[p], end = pbeg[p + 1];
  const uint32_t same = psame ? psame[p] : 0u;
  const int32_t* __restrict__ pc = pcols + (size_t)p * kSlots;
  if (count <= narrow) {
    panel_pairs<EPI>(beg, end, pc, outs, w, bias_vec, bias_s, clip, yin, ldin, yout, ldout,
                     count, map, flags, zc, kcap, mcent, eps_post);
    return;
  }
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  int64_t run0, run1;
  tile_run(count, run0, run1);
  for (int64_t base = run0 + wv * kWave; base < run1; base += kBlock) {
    const int64_t pos = base + lane;
    const bool valid = pos < count;
    float x[kSlots];
    bool nz = false;
    if (valid) {
      const int64_t src = map ? (int64_t)map[pos] : pos;
#pragma unroll
      for (int r = 0; r < kSlots; ++r) {
        x[r] = yin[(int64_t)pc[r] * ldin + src];
        nz |= (x[r] != 0.f);
      }
    } else {
#pragma unroll
      for (int r = 0; r < kSlots; ++r) x[r] = 0.f;
    }
    const bool wave_nz = __ballot(nz) != 0ull;
    int32_t mi = 0;
    if (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) mi = valid ? mcent[pos] : 0;
    bool flag = false;
    float acc = 0.f;
    for (int j = beg; j < end; ++j) {
      // a row repeating the previous row's weights has its chain (same inputs, same order)
      if (!((same >> (j - beg)) & 1u)) {
        const float* __restrict__ wj = w + (size_t)j * kSlots;
        acc = 0.f;
        if (wave_nz) {
#pragma unroll
          for (int r = 0; r < kSlots; ++r) acc = __builtin_fmaf(wj[r], x[r], acc);
        }
      }
      const int o = outs[j];
      const float b = bias_vec ? bias_vec[o] : bias_s;
      const float y = epilogue<EPI>(acc, b, clip, valid, o, pos, mi, zc, kcap, eps_post, flag);
      if (valid) yout[(int64_t)o * ldout + pos] = y;
    }
    if (flags && valid && flag) flags[pos] = 1;
  }
}

// Plain panel layer (EPI_PLAIN, vector chains) with TPW neighbouring 64-column tiles per
// wavefront and iteration: a workgroup takes runs of 256 * TPW columns, wavefront w the
// tiles base + 64 (TPW w + q), and each wavefront issues the gathers of all its tiles before
// the chains.  When the column list is shifted, neighbouring tiles share a 128-byte line;
// within one wavefront both requests reach the cache together and the line is fetched once,
// while tiles of different wavefronts request it separately and a third of such lines were
// fetched twice (L2 counters on the 16384-neuron net).  Same results as panel_spmm_kernel
// <EPI_PLAIN>; more registers (102 at TPW 2, 158 at 4; 64 at one tile) and fewer wavefronts
// per SIMD.  Pre-threshold stage, same plan and placement: 103.1 ms (one tile) against
// 102.1 (2) and 101.9 (4) at N=65536, 27.44 against 27.25 (2) and 27.35 (4) at N=16384.
template <int TPW>
__global__ __launch_bounds__(kBlock, (TPW == 2 ? 4 : 2)) void panel_spmm_tiles_kernel(
    const int32_t* __restrict__ pcols, const int32_t* __restrict__ pbeg,
    const int32_t* __restrict__ outs, const float* __restrict__ w,
    const float* __restrict__ bias_vec, float bias_s, float clip, const float* __restrict__ yin,
    int64_t ldin, float* __restrict__ yout, int64_t ldout, const int32_t* __restrict__ count_ptr,
    const int32_t* __restrict__ map, int32_t* __restrict__ flags, int32_t narrow,
    const uint32_t* __restrict__ psame) {
  constexpr int64_t kSpan = (int64_t)kBlock * TPW;
  const int p = blockIdx.x;
  const int64_t count = *count_ptr;
  const int beg = pbeg[p], end = pbeg[p + 1];
  const uint32_t same = psame ? psame[p] : 0u;
  const int32_t* __restrict__ pc = pcols + (size_t)p * kSlots;
  if (count <= narrow) {
    panel_pairs<EPI_PLAIN>(beg, end, pc, outs, w, bias_vec, bias_s, clip, yin, ldin, yout,
                           ldout, count, map, flags, nullptr, 0, nullptr, 0.f);
    return;
  }
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  // this workgroup row's run of kSpan-column units, [r0, r1)
  const int64_t units = (count + kSpan - 1) / kSpan;
  const int64_t per = (units + gridDim.y - 1) / gridDim.y;
  const int64_t u0 = min(units, (int64_t)blockIdx.y * per);
  const int64_t r0 = u0 * kSpan, r1 = min(count, (u0 + per) * kSpan);
  for (int64_t base = r0 + (int64_t)wv * kWave * TPW; base < r1; base += kSpan) {
    int64_t ps[TPW];
    bool vs[TPW];
    float x[TPW][kSlots];
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      ps[q] = base + kWave * q + lane;
      vs[q] = ps[q] < r1;
      // positions past the run read column 0 (count > 0 here) and are zeroed below
      const int64_t src = vs[q] ? (map ? (int64_t)map[ps[q]] : ps[q]) : 0;
#pragma unroll
      for (int r = 0; r < kSlots; ++r) x[q][r] = yin[(int64_t)pc[r] * ldin + src];
    }
    bool nz = false;
#pragma unroll
    for (int q = 0; q < TPW; ++q)
#pragma unroll
      for (int r = 0; r < kSlots; ++r) {
        x[q][r] = vs[q] ? x[q][r] : 0.f;
        nz |= x[q][r] != 0.f;
      }
    const bool wave_nz = __ballot(nz) != 0ull;
    bool fl[TPW];
    float acc[TPW];
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      fl[q] = false;
      acc[q] = 0.f;
    }
    for (int j = beg; j < end; ++j) {
      // a row repeating the previous row's weights has its chains (same inputs, same order)
      if (!((same >> (j - beg)) & 1u)) {
        const float* __restrict__ wj = w + (size_t)j * kSlots;
#pragma unroll
        for (int q = 0; q < TPW; ++q) acc[q] = 0.f;
        if (wave_nz) {
#pragma unroll
          for (int r = 0; r < kSlots; ++r)
#pragma unroll
            for (int q = 0; q < TPW; ++q) acc[q] = __builtin_fmaf(wj[r], x[q][r], acc[q]);
        }
      }
      const int o = outs[j];
      const float b = bias_vec ? bias_vec[o] : bias_s;
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        const float y = epilogue<EPI_PLAIN>(acc[q], b, clip, vs[q], o, ps[q], 0, nullptr, 0, 0.f,
                                            fl[q]);
        if (vs[q]) yout[(int64_t)o * ldout + ps[q]] = y;
      }
    }
#pragma unroll
    for (int q = 0; q < TPW; ++q)
      if (flags && vs[q] && fl[q]) flags[ps[q]] = 1;
  }
}

// ---- plain panel layer on the matrix cores ----------------------------------------------
// Same contract as panel_spmm_kernel<EPI_PLAIN> for row-major panel weights.  The workgroup
// stages its panel's weights in LDS (rows past the panel's row count as zeros), then each
// wavefront owns 64 columns per tile: it gathers the 32 inputs of its columns into registers
// and runs slot r as one v_mfma_f32_32x32x1_2b_f32 step, A = the 32 row weights of slot r
// (both 32-lane halves), B = the 64 columns' input of slot r (block 0 = columns 0..31,
// block 1 = 32..63).  With K = 1 every element of a step is one fused multiply-add, so the
// 32 steps in slot order are exactly the vector kernel's chain
// acc = fma(w[31], x[31], ... fma(w[0], x[0], 0)) (src/SDGC/BF/kernel.hpp:57-88).  The
// matrix cores take the 32 multiply-adds per output off the vector ALU, which otherwise
// holds the layer below the memory rate; 96 registers give 5 wavefronts per SIMD.
// D element (block b, row i, column j) sits in lane 32 (i / 4 % 2) + j, register
// 16 b + 4 (i / 8) + i % 4.  Up to `narrow` columns the (row, column) pair path runs.
__global__ __launch_bounds__(kBlock, 5) void panel_plain_mfma_kernel(
    const int32_t* __restrict__ pcols, const int32_t* __restrict__ pbeg,
    const int32_t* __restrict__ outs, const float* __restrict__ w,
    const float* __restrict__ bias_vec, float bias_s, float clip, const float* __restrict__ yin,
    int64_t ldin, float* __restrict__ yout, int64_t ldout, const int32_t* __restrict__ count_ptr,
    const int32_t* __restrict__ map, int32_t* __restrict__ flags, int32_t narrow) {
  __shared__ float w_s[kMaxPanelRows * (kSlots + 1)];  // row i, slot r at i * 33 + r
  __shared__ int32_t o_s[kMaxPanelRows];
  __shared__ float b_s[kMaxPanelRows];
  const int p = blockIdx.x;
  const int64_t count = *count_ptr;
  const int beg = pbeg[p], end = pbeg[p + 1];
  const int32_t* __restrict__ pc = pcols + (size_t)p * kSlots;
  if (count <= narrow) {
    panel_pairs<EPI_PLAIN>(beg, end, pc, outs, w, bias_vec, bias_s, clip, yin, ldin, yout,
                           ldout, count, map, flags, nullptr, 0, nullptr, 0.f);
    return;
  }
  const int G = end - beg;
  for (int q = threadIdx.x; q < kMaxPanelRows * kSlots; q += kBlock) {
    const int i = q / kSlots, r = q % kSlots;
    w_s[i * (kSlots + 1) + r] = i < G ? w[(size_t)(beg + i) * kSlots + r] : 0.f;
  }
  if (threadIdx.x < kMaxPanelRows) {
    const int i = threadIdx.x;
    const int o = i < G ? outs[beg + i] : 0;
    o_s[i] = o;
    b_s[i] = bias_vec ? bias_vec[o] : bias_s;
  }
  __syncthreads();
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int half = lane >> 5, jl = lane & 31;
  int64_t run0, run1;
  tile_run(count, run0, run1);
  for (int64_t base = run0 + wv * kWave; base < run1; base += kBlock) {
    const int64_t pos = base + lane;
    const bool valid = pos < count;
    // columns past the count read column 0 (count > 0 here) and are zeroed
    const int64_t src = valid ? (map ? (int64_t)map[pos] : pos) : 0;
    float x[kSlots];
#pragma unroll
    for (int r = 0; r < kSlots; ++r) x[r] = yin[(int64_t)pc[r] * ldin + src];
    bool nz = false;
#pragma unroll
    for (int r = 0; r < kSlots; ++r) {
      x[r] = valid ? x[r] : 0.f;
      nz |= x[r] != 0.f;
    }
    floatx32 acc;
#pragma unroll
    for (int q = 0; q < 32; ++q) acc[q] = 0.f;
    // all 64 x 32 inputs zero: acc = +0 exactly, what the steps would give
    if (__ballot(nz) != 0ull) {
#pragma unroll
      for (int r = 0; r < kSlots; ++r)
        acc = __builtin_amdgcn_mfma_f32_32x32x1f32(w_s[jl * (kSlots + 1) + r], x[r], acc, 0, 0,
                                                   0);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int64_t pb = base + 32 * b + jl;
      const bool vb = pb < count;
      bool flag = false;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int i = 8 * (q >> 2) + 4 * half + (q & 3);
        if (i < G) {
          const float y = act_clip(acc[16 * b + q] + b_s[i], clip);
          flag |= y > 0.f;
          if (vb) yout[(int64_t)o_s[i] * ldout + pb] = y;
        }
      }
      // the column's rows are split over lanes j and j + 32
      const uint64_t m = __ballot(flag);
      const bool any = ((m >> jl) & 1ull) | ((m >> (jl + 32)) & 1ull);
      if (flags && half == 0 && vb && any) flags[pb] = 1;
    }
  }
}

// ---- post-threshold centroid rows --------------------------------------------------------
// Centroid rows of one post layer: zc = W yhat_c + b and yhat_c' = f(zc) for the k centroid
// columns (post_minus, src/SDGC/SNICIT/kernel.hpp:115-121).  One block per panel stages the
// panel's weights and its 32 inputs of every centroid in LDS, then thread (row, centroid)
// runs the fixed-order chain.
__global__ __launch_bounds__(kBlock) void panel_centroid_kernel(
    const int32_t* __restrict__ pcols, const int32_t* __restrict__ pbeg,
    const int32_t* __restrict__ outs, const float* __restrict__ w,
    const float* __restrict__ bias_vec, float bias_s, float clip, const float* __restrict__ cin,
    float* __restrict__ cout, float* __restrict__ zc, int32_t kcap,
    const int32_t* __restrict__ k_ptr) {
  __shared__ float4 w_s[kMaxPanelRows * kSlots / 4];
  __shared__ float x_s[kSlots * (kMaxCentroidsExact + 1)];
  const int p = blockIdx.x;
  const int k = *k_ptr;
  if (k == 0) return;
  const int beg = pbeg[p], G = pbeg[p + 1] - beg;
  const int32_t* __restrict__ pc = pcols + (size_t)p * kSlots;
  const float4* __restrict__ w4 = reinterpret_cast<const float4*>(w + (size_t)beg * kSlots);
  for (int q = threadIdx.x; q < G * kSlots / 4; q += kBlock) w_s[q] = w4[q];
  for (int q = threadIdx.x; q < kSlots * k; q += kBlock) {
    const int r = q / k, c = q % k;
    x_s[r * (kMaxCentroidsExact + 1) + c] = cin[(int64_t)pc[r] * kcap + c];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < G * k; q += kBlock) {
    const int jj = q / k, c = q % k;
    float acc = 0.f;
#pragma unroll
    for (int r4 = 0; r4 < kSlots / 4; ++r4) {
      const float4 wv = w_s[jj * (kSlots / 4) + r4];
      acc = __builtin_fmaf(wv.x, x_s[(4 * r4 + 0) * (kMaxCentroidsExact + 1) + c], acc);
      acc = __builtin_fmaf(wv.y, x_s[(4 * r4 + 1) * (kMaxCentroidsExact + 1) + c], acc);
      acc = __builtin_fmaf(wv.z, x_s[(4 * r4 + 2) * (kMaxCentroidsExact + 1) + c], acc);
      acc = __builtin_fmaf(wv.w, x_s[(4 * r4 + 3) * (kMaxCentroidsExact + 1) + c], acc);
    }
    const int o = outs[beg + jj];
    const float z = acc + (bias_vec ? bias_vec[o] : bias_s);
    zc[(int64_t)o * kcap + c] = z;
    cout[(int64_t)o * kcap + c] = act_clip(z, clip);
  }
}

// ---- post-threshold layer -------------------------------------------------------------
// One launch per post layer (src/SDGC/SNICIT/kernel.hpp:71-136, post_spMM + post_minus):
// the centroid rows zc = W yhat_c + b of every panel (workgroup row 0 also writes
// yhat_c' = f(zc)), then the delta rows of the launch's columns,
// y = f(zc[m] + W yhat_i) - f(zc[m]).  Same fused-multiply-add chain as panel_spmm_kernel.
// psame (nullable) marks rows whose weights equal the previous row's: their chain is the
// previous row's, bit for bit, so it is computed once (SDGC panels: one chain for 32 rows).
// Two forms, chosen per panel:
//  * at most two distinct weight rows and more than `narrow` columns: every wavefront works
//    alone, with no LDS and no barrier.  Lanes hold the distinct rows' weights (lanes 0-31
//    and 32-63) and the rows' output ids and biases; lane d*k + c runs the centroid chain of
//    distinct row d and centroid c, and each column takes its centroid's value with a
//    shuffle.  The wavefront's column gathers and its centroid loads are in flight together.
//  * otherwise the workgroup stages the panel's weights, its 32 inputs of every centroid
//    and the centroid rows in LDS; with at most `narrow` columns one thread per (row,
//    column) pair, else each wavefront owns 64 columns and every row of the panel.
// Compact post layout (ldin_cnt non-null, compact_out): the delta buffers' row stride is
// compact_ld of the positions they hold, read from the device counters, so a layer's delta
// rows fill one dense region; nlist_next (nullable) receives the list length for the next
// layer when no compaction runs in between.  ccompact: the centroid columns cin/cout have
// the stride compact_kst(k, kcap) instead of kcap (fewer bytes per neuron when k is small).

// Panel chain record of the wave-local form: centroid chain (before the bias) of distinct row
// d and centroid c at d * 32 + c, one value per lane of a wavefront (d < 2, c < 32).
constexpr int kChainRec = kPostChainRec;
static_assert(kChainRec == 2 * kMaxCentroidsExact, "one chain per lane of a wavefront");

// Second distinct row of a panel with at most two: the first row past row 0 whose weights
// differ from the previous row's (G when every row has row 0's weights).
__device__ __forceinline__ int second_row(const uint32_t* __restrict__ psame, int p, int G) {
  const uint32_t same = psame ? psame[p] : 0u;
  const uint32_t diff = ~same & (G >= 32 ? 0xfffffffeu : ((1u << G) - 1u) & ~1u);
  return diff ? __ffs(diff) - 1 : G;
}

// Centroid chains of one panel on one wavefront: lane d * 32 + c runs the chain of distinct
// row d (row 0, or row r1) and centroid c over the panel's 32 inputs of centroid column c
// (cin, stride kst), in slot order; lanes past the panel's rows or centroids run a clamped
// copy whose value is never taken.  CHUNK inputs are loaded at a time (8 where the caller
// holds the tile's gathers meanwhile, 32 otherwise: one round trip).
template <int CHUNK>
__device__ __forceinline__ float panel_chain(int lane, int G, int r1, int beg,
                                             const int32_t* __restrict__ pc,
                                             const float* __restrict__ w,
                                             const float* __restrict__ cin, int kst, int k) {
  const int d = (lane >> 5) && r1 < G ? 1 : 0;
  const int c = min(lane & 31, (k > 0 ? k : 1) - 1);
  const float* __restrict__ wr = w + (size_t)(beg + (d ? r1 : 0)) * kSlots;
  float z = 0.f;
#pragma unroll
  for (int r0 = 0; r0 < kSlots; r0 += CHUNK) {
    float xc[CHUNK], wc[CHUNK];
#pragma unroll
    for (int r = 0; r < CHUNK; ++r) {
      xc[r] = cin[(int64_t)pc[r0 + r] * kst + c];
      wc[r] = wr[r0 + r];
    }
#pragma unroll
    for (int r = 0; r < CHUNK; ++r) z = __builtin_fmaf(wc[r], xc[r], z);
  }
  return z;
}

// Centroid rows of one panel from its chain record (lane d * 32 + c): cout[o][c] =
// f(chain(d(o), c) + b(o)) for the panel's rows o, d(o) = 1 from row r1 on.  Lane jj < G
// holds row jj's output id (ol) and bias (bl), so the stores need no loads.
__device__ __forceinline__ void panel_centroid_rows(int lane, int G, int r1, int ol, float bl,
                                                    float clip, float zch,
                                                    float* __restrict__ cout, int kst, int k) {
  const int d = lane >> 5, c = lane & 31;
  const bool active = c < k && !(d == 1 && r1 >= G);
  const int jb = d ? r1 : 0, je = d ? G : r1;
  for (int jj = 0; jj < G; ++jj) {
    const int o = __builtin_amdgcn_readlane(ol, jj);
    const float b = readlane_f(bl, jj);
    if (active && jj >= jb && jj < je) cout[(int64_t)o * kst + c] = act_clip(zch + b, clip);
  }
}

// Wave-local form (see above).  G rows with one distinct weight row, or two (TWO) where the
// second starts at row r1.  The wavefront's first column gathers are issued before anything
// else, so every wavefront has loads in flight from its start.  Its panel's chain record
// comes from zrec (precomputed by post_chain_kernel) or is computed here.  Each tile takes
// its columns' centroid values with one shuffle per distinct row; with a scalar bias every
// row of a distinct row's class has the same output.
template <bool LOSSY, bool TWO>
__device__ __forceinline__ void post_wave_form(
    int G, int r1, int beg, const int32_t* __restrict__ pc, const int32_t* __restrict__ outs,
    const float* __restrict__ w, const float* __restrict__ bias_vec, float bias_s, float clip,
    const float* __restrict__ zrec, const float* __restrict__ cin, float* __restrict__ cout,
    int kst, int k, const float* __restrict__ din, int64_t ldin, float* __restrict__ dout,
    int64_t ldout, int64_t count, const int32_t* __restrict__ map,
    const int32_t* __restrict__ mcent, int32_t* __restrict__ flags, float eps_post) {
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int64_t step = (int64_t)gridDim.y * kBlock;
  int64_t base = (int64_t)blockIdx.y * kBlock + wv * kWave;
  // The tile at `base`: columns past count load list position 0 (count > 0 whenever a tile
  // runs) and store nothing; every load is unconditional, so the gathers and the centroid
  // index are in flight together.
  float x[kSlots];
  int32_t mi = 0;
  auto load_tile = [&](int64_t b) {
    const int64_t pos = b + lane;
    const uint32_t col = (uint32_t)(pos < count ? pos : 0);
    const uint32_t src = map ? (uint32_t)map[col] : col;
#pragma unroll
    for (int r = 0; r < kSlots; ++r) x[r] = din[(int64_t)pc[r] * ldin + src];
    mi = mcent[col];
  };
  if (base < count) load_tile(base);
  // lanes 0-31: weights of row 0 (the first distinct row); lanes 32-63: row r1 (TWO; row 0
  // again when the panel has one distinct row, r1 = G)
  const float wl = w[(size_t)(beg + (TWO && lane >= 32 && r1 < G ? r1 : 0)) * kSlots + (lane & 31)];
  const int ol = outs[beg + (lane < G ? lane : 0)];
  const float bl = bias_vec ? bias_vec[ol] : bias_s;
  float zch;
  if (zrec) {
    zch = zrec[lane];
  } else {
    zch = panel_chain<8>(lane, G, r1, beg, pc, w, cin, kst, k);
    if (blockIdx.y == 0 && wv == 0 && cout)
      panel_centroid_rows(lane, G, r1, ol, bl, clip, zch, cout, kst, k);
  }
  auto residue = [&](float acc, float z) {
    float y = act_clip(z + acc, clip) - act_clip(z, clip);
    if (LOSSY) y = fabsf(y) > eps_post ? y : 0.f;
    return y;
  };
  while (base < count) {
    const int64_t pos = base + lane;
    const bool valid = pos < count;
    bool nz = false;
#pragma unroll
    for (int r = 0; r < kSlots; ++r) nz |= (x[r] != 0.f);
    // a wavefront whose inputs are all zero keeps acc = +0, what the chain gives
    float acc0 = 0.f, acc1 = 0.f;
    if (__ballot(nz) != 0ull) {
#pragma unroll
      for (int r = 0; r < kSlots; ++r) acc0 = __builtin_fmaf(readlane_f(wl, r), x[r], acc0);
      if (TWO) {
#pragma unroll
        for (int r = 0; r < kSlots; ++r) acc1 = __builtin_fmaf(readlane_f(wl, 32 + r), x[r], acc1);
      }
    }
    // centroid value of each column's centroid, per distinct row (before the bias)
    const float zs0 = __shfl(zch, mi);
    const float zs1 = TWO ? __shfl(zch, 32 + mi) : 0.f;
    bool flag = false;
    if (valid) {
      float* __restrict__ ocol = dout + pos;
      if (!bias_vec) {
        const float y0 = residue(acc0, zs0 + bias_s);
        const float y1 = TWO ? residue(acc1, zs1 + bias_s) : 0.f;
        flag = y0 != 0.f || (TWO && y1 != 0.f);
        if (G == kMaxPanelRows) {
#pragma unroll
          for (int jj = 0; jj < kMaxPanelRows; ++jj)
            ocol[(int64_t)__builtin_amdgcn_readlane(ol, jj) * ldout] = (TWO && jj >= r1) ? y1 : y0;
        } else {
          for (int jj = 0; jj < G; ++jj)
            ocol[(int64_t)__builtin_amdgcn_readlane(ol, jj) * ldout] = (TWO && jj >= r1) ? y1 : y0;
        }
      } else {
        for (int jj = 0; jj < G; ++jj) {
          const bool second = TWO && jj >= r1;
          const float y = residue(second ? acc1 : acc0, (second ? zs1 : zs0) + readlane_f(bl, jj));
          flag |= (y != 0.f);
          ocol[(int64_t)__builtin_amdgcn_readlane(ol, jj) * ldout] = y;
        }
      }
      if (flag) flags[pos] = 1;
    }
    base += step;
    if (base < count) load_tile(base);
  }
}

// Chain record and centroid rows of panel p of a post layer on one wavefront (see
// panel_chain): zchain[p] = the panel's chains, cout = f(chain + b) from cin (stride kst).
__device__ __forceinline__ void chain_panel(int64_t p, const int32_t* __restrict__ pcols,
                                            const int32_t* __restrict__ pbeg,
                                            const int32_t* __restrict__ outs,
                                            const float* __restrict__ w,
                                            const uint32_t* __restrict__ psame,
                                            const float* __restrict__ bias_vec, float bias_s,
                                            float clip, const float* __restrict__ cin,
                                            float* __restrict__ cout, int kst, int k,
                                            float* __restrict__ zchain) {
  const int lane = lane_id();
  const int beg = pbeg[p], G = pbeg[p + 1] - beg;
  const int r1 = second_row(psame, (int)p, G);
  const int ol = outs[beg + min(lane & 31, G - 1)];
  const float bl = bias_vec ? bias_vec[ol] : bias_s;
  const float z = panel_chain<kSlots>(lane, G, r1, beg, pcols + (size_t)p * kSlots, w, cin, kst, k);
  zchain[p * kChainRec + lane] = z;
  panel_centroid_rows(lane, G, r1, ol, bl, clip, z, cout, kst, k);
}

// zchain (nullable): the layer's chain records, kChainRec floats per panel (from the launch
// of the layer before, or post_chain_kernel); null: each wavefront computes its panel's
// chains from cin, and the lead wavefront writes cout.
// nx (nx.pbeg non-null): the next post layer's chain records and centroid rows, computed
// here by a fifth wavefront (workgroups of kBlock + 64 threads) in workgroup row 0 of each
// panel, beside the four that stream the delta rows (grid.x = the larger of the two layers'
// panel counts; P = this layer's).  They depend on the centroid columns only and the next
// launch needs them, so no layer after the first needs a launch of its own for them.
template <bool LOSSY, bool TWO>
__global__ __launch_bounds__(kBlock + kWave, 8) void panel_post_wave_kernel(
    const int32_t* __restrict__ pcols, const int32_t* __restrict__ pbeg,
    const int32_t* __restrict__ outs, const float* __restrict__ w,
    const float* __restrict__ bias_vec, float bias_s, float clip,
    const float* __restrict__ zchain, const float* __restrict__ cin, float* __restrict__ cout,
    int32_t kcap, const int32_t* __restrict__ k_ptr, const float* __restrict__ din,
    int64_t ldin_arg, const int32_t* __restrict__ ldin_cnt, float* __restrict__ dout,
    int64_t ldout_arg, int32_t compact_out, const int32_t* __restrict__ count_ptr,
    int32_t* __restrict__ nlist_next, const int32_t* __restrict__ map,
    const int32_t* __restrict__ mcent, int32_t* __restrict__ flags, float eps_post,
    const uint32_t* __restrict__ psame, int32_t ccompact, int64_t P, NextChain nx) {
  const int p = blockIdx.x;
  const int k = *k_ptr;
  const int kst = ccompact ? compact_kst(k, kcap) : kcap;
  if (threadIdx.x >= kBlock) {  // the fifth wavefront: next layer's chains only
    if (nx.pbeg && blockIdx.y == 0 && p < nx.P)
      chain_panel(p, nx.pcols, nx.pbeg, nx.outs, nx.w, nx.psame, nx.bias_vec, nx.bias_s, clip,
                  nx.cin, nx.cout, kst, k, nx.zchain);
    return;
  }
  if (p >= P) return;
  const int64_t count = *count_ptr;
  const int beg = pbeg[p], G = pbeg[p + 1] - beg;
  const int64_t ldin = ldin_cnt ? compact_ld(*ldin_cnt) : ldin_arg;
  const int64_t ldout = compact_out ? compact_ld(count) : ldout_arg;
  if (nlist_next && p == 0 && blockIdx.y == 0 && threadIdx.x == 0) *nlist_next = (int32_t)count;
  // without chain records workgroup row 0 always runs: it writes the centroid rows
  if (blockIdx.y != 0 && (int64_t)blockIdx.y * kBlock >= count) return;
  if (zchain && (int64_t)blockIdx.y * kBlock >= count) return;
  const int r1 = TWO ? second_row(psame, p, G) : G;
  post_wave_form<LOSSY, TWO>(G, r1, beg, pcols + (size_t)p * kSlots, outs, w, bias_vec, bias_s,
                             clip, zchain ? zchain + (size_t)p * kChainRec : nullptr, cin, cout,
                             kst, k, din, ldin, dout, ldout, count, map, mcent, flags, eps_post);
}

// Centroid rows of one post layer whose panels have at most two distinct weight rows, apart
// from the delta rows: one wavefront per panel computes its chain record (zchain, kChainRec
// floats per panel) and the centroid rows cout = f(chain + b) from cin (stride
// compact_kst(k, kcap) when ccompact, else kcap).
__global__ __launch_bounds__(kBlock) void post_chain_kernel(
    const int32_t* __restrict__ pcols, const int32_t* __restrict__ pbeg,
    const int32_t* __restrict__ outs, const float* __restrict__ w,
    const uint32_t* __restrict__ psame, const float* __restrict__ bias_vec, float bias_s,
    float clip, const float* __restrict__ cin, float* __restrict__ cout, int32_t kcap,
    const int32_t* __restrict__ k_ptr, int32_t ccompact, float* __restrict__ zchain, int64_t P) {
  const int64_t p = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (p >= P) return;
  const int k = *k_ptr;
  const int kst = ccompact ? compact_kst(k, kcap) : kcap;
  chain_panel(p, pcols, pbeg, outs, w, psame, bias_vec, bias_s, clip, cin, cout, kst, k, zchain);
}

template <bool LOSSY>
__global__ __launch_bounds__(kBlock) void panel_post_kernel(
    const int32_t* __restrict__ pcols, const int32_t* __restrict__ pbeg,
    const int32_t* __restrict__ outs, const float* __restrict__ w,
    const float* __restrict__ bias_vec, float bias_s, float clip, const float* __restrict__ cin,
    float* __restrict__ cout, int32_t kcap, const int32_t* __restrict__ k_ptr,
    const float* __restrict__ din, int64_t ldin_arg, const int32_t* __restrict__ ldin_cnt,
    float* __restrict__ dout, int64_t ldout_arg, int32_t compact_out,
    const int32_t* __restrict__ count_ptr, int32_t* __restrict__ nlist_next,
    const int32_t* __restrict__ map, const int32_t* __restrict__ mcent,
    int32_t* __restrict__ flags, float eps_post, int32_t narrow,
    const uint32_t* __restrict__ psame, int32_t ccompact) {
  constexpr int kC = kMaxCentroidsExact + 1;  // padded centroid stride in LDS
  __shared__ float4 w_s[kMaxPanelRows * kSlots / 4];
  __shared__ float x_s[kSlots * kC];
  __shared__ float zc_s[kMaxPanelRows * kC];
  __shared__ int32_t o_s[kMaxPanelRows];
  const int p = blockIdx.x;
  const int64_t count = *count_ptr;
  const int k = *k_ptr;
  const int beg = pbeg[p], G = pbeg[p + 1] - beg;
  const uint32_t same = psame ? psame[p] : 0u;
  const int64_t ldin = ldin_cnt ? compact_ld(*ldin_cnt) : ldin_arg;
  const int64_t ldout = compact_out ? compact_ld(count) : ldout_arg;
  const int kst = ccompact ? compact_kst(k, kcap) : kcap;
  if (nlist_next && p == 0 && blockIdx.y == 0 && threadIdx.x == 0) *nlist_next = (int32_t)count;
  // narrow: every workgroup of the panel takes (row, column) pairs; wide: 256 columns each.
  // Workgroup row 0 always runs: it writes the centroid rows, delta rows or not.
  if (blockIdx.y != 0 && (count <= narrow ? (int64_t)blockIdx.y * kBlock >= (int64_t)G * count
                                          : (int64_t)blockIdx.y * kBlock >= count))
    return;
  const int32_t* __restrict__ pc = pcols + (size_t)p * kSlots;
  {
    const float4* __restrict__ w4 = reinterpret_cast<const float4*>(w + (size_t)beg * kSlots);
    for (int q = threadIdx.x; q < G * kSlots / 4; q += kBlock) w_s[q] = w4[q];
    if (threadIdx.x < G) o_s[threadIdx.x] = outs[beg + threadIdx.x];
    for (int q = threadIdx.x; q < kSlots * k; q += kBlock) {
      const int r = q / k, c = q % k;
      x_s[r * kC + c] = cin[(int64_t)pc[r] * kst + c];
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < G * k; q += kBlock) {
    const int jj = q / k, c = q % k;
    float acc = 0.f;
#pragma unroll
    for (int r4 = 0; r4 < kSlots / 4; ++r4) {
      const float4 wv = w_s[jj * (kSlots / 4) + r4];
      acc = __builtin_fmaf(wv.x, x_s[(4 * r4 + 0) * kC + c], acc);
      acc = __builtin_fmaf(wv.y, x_s[(4 * r4 + 1) * kC + c], acc);
      acc = __builtin_fmaf(wv.z, x_s[(4 * r4 + 2) * kC + c], acc);
      acc = __builtin_fmaf(wv.w, x_s[(4 * r4 + 3) * kC + c], acc);
    }
    const int o = o_s[jj];
    const float z = acc + (bias_vec ? bias_vec[o] : bias_s);
    zc_s[jj * kC + c] = z;
    if (blockIdx.y == 0 && cout) cout[(int64_t)o * kst + c] = act_clip(z, clip);
  }
  __syncthreads();
  auto residue = [&](float acc, float z) {
    float y = act_clip(z + acc, clip) - act_clip(z, clip);
    if (LOSSY) y = fabsf(y) > eps_post ? y : 0.f;
    return y;
  };
  if (count <= narrow) {
    const int64_t pairs = (int64_t)G * count;
    for (int64_t q = (int64_t)blockIdx.y * kBlock + threadIdx.x; q < pairs;
         q += (int64_t)gridDim.y * kBlock) {
      const int jj = (int)(q % G);
      const int64_t pos = q / G;
      const int64_t src = map ? (int64_t)map[pos] : pos;
      float xn[kSlots];
#pragma unroll
      for (int r = 0; r < kSlots; ++r) xn[r] = din[(int64_t)pc[r] * ldin + src];
      float acc = 0.f;
#pragma unroll
      for (int r4 = 0; r4 < kSlots / 4; ++r4) {
        const float4 wv = w_s[jj * (kSlots / 4) + r4];
        acc = __builtin_fmaf(wv.x, xn[4 * r4 + 0], acc);
        acc = __builtin_fmaf(wv.y, xn[4 * r4 + 1], acc);
        acc = __builtin_fmaf(wv.z, xn[4 * r4 + 2], acc);
        acc = __builtin_fmaf(wv.w, xn[4 * r4 + 3], acc);
      }
      const float y = residue(acc, zc_s[jj * kC + mcent[pos]]);
      dout[(int64_t)o_s[jj] * ldout + pos] = y;
      if (y != 0.f) flags[pos] = 1;
    }
    return;
  }
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  int64_t run0, run1;
  tile_run(count, run0, run1);
  for (int64_t base = run0 + wv * kWave; base < run1; base += kBlock) {
    const int64_t pos = base + lane;
    const bool valid = pos < count;
    float x[kSlots];
    bool nz = false;
    int32_t mi = 0;
    if (valid) {
      const int64_t src = map ? (int64_t)map[pos] : pos;
      mi = mcent[pos];
#pragma unroll
      for (int r = 0; r < kSlots; ++r) {
        x[r] = din[(int64_t)pc[r] * ldin + src];
        nz |= (x[r] != 0.f);
      }
    } else {
#pragma unroll
      for (int r = 0; r < kSlots; ++r) x[r] = 0.f;
    }
    const bool wave_nz = __ballot(nz) != 0ull;
    bool flag = false;
    float acc = 0.f;
    for (int jj = 0; jj < G; ++jj) {
      if (!((same >> jj) & 1u)) {
        acc = 0.f;
        if (wave_nz) {
#pragma unroll
          for (int r4 = 0; r4 < kSlots / 4; ++r4) {
            const float4 wv4 = w_s[jj * (kSlots / 4) + r4];
            acc = __builtin_fmaf(wv4.x, x[4 * r4 + 0], acc);
            acc = __builtin_fmaf(wv4.y, x[4 * r4 + 1], acc);
            acc = __builtin_fmaf(wv4.z, x[4 * r4 + 2], acc);
            acc = __builtin_fmaf(wv4.w, x[4 * r4 + 3], acc);
          }
        }
      }
      const float y = residue(acc, zc_s[jj * kC + mi]);
      flag |= (y != 0.f);
      if (valid) dout[(int64_t)o_s[jj] * ldout + pos] = y;
    }
    if (valid && flag) flags[pos] = 1;
  }
}

// Centroid columns between the kcap stride and the compact stride compact_kst(k, kcap).
__global__ __launch_bounds__(256) void centroid_restride_kernel(
    const float* __restrict__ src, int32_t src_compact, float* __restrict__ dst,
    int32_t dst_compact, int32_t kcap, const int32_t* __restrict__ k_ptr, int64_t N) {
  const int k = *k_ptr;
  const int64_t ks = compact_kst(k, kcap);
  const int64_t ss = src_compact ? ks : kcap, ds = dst_compact ? ks : kcap;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N * k;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = q / k;
    const int c = (int)(q - n * k);
    dst[n * ds + c] = src[n * ss + c];
  }
}

// ---- panel SpMM on the matrix cores ------------------------------------------------
// Same contract as panel_spmm_kernel for layers stored slot-major (panels of <= 32 rows).
// A wavefront owns 64 consecutive columns; each slot r is one v_mfma_f32_32x32x1_2b_f32:
// A = the panel's 32 row weights of slot r (both 32-lane halves), B = the 64 columns' input
// of slot r (block 0 = columns 0..31, block 1 = 32..63), D += A x B.  With K = 1 every
// element is one fused multiply-add, so the 32 steps in slot order are exactly the chain
// acc = fma(w[31], x[31], ... fma(w[0], x[0], 0)) of the vector kernel and the oracle.
// D element (block b, row i, column j) sits in lane 32*(i/4 % 2) + j, register
// 16 b + 4 (i / 8) + i % 4: each store writes two 128-byte runs (rows i, i+4 of 32 columns).

template <int EPI>
__global__ __launch_bounds__(kBlock, 2) void panel_mfma_kernel(
    const int32_t* __restrict__ pcols, const int32_t* __restrict__ pbeg,
    const int32_t* __restrict__ outs, const float* __restrict__ wslot,
    const float* __restrict__ bias_vec, float bias_s, float clip, const float* __restrict__ yin,
    int64_t ldin, float* __restrict__ yout, int64_t ldout, const int32_t* __restrict__ count_ptr,
    const int32_t* __restrict__ map, int32_t* __restrict__ flags, float* __restrict__ zc,
    int32_t kcap, const int32_t* __restrict__ mcent, float eps_post) {
  const int p = blockIdx.x;
  const int64_t count = *count_ptr;
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int64_t first = (int64_t)blockIdx.y * kBlock + wv * kWave;
  if (first >= count) return;
  const int beg = pbeg[p], G = pbeg[p + 1] - beg;
  const int32_t* __restrict__ pc = pcols + (size_t)p * kSlots;
  const int half = lane >> 5, j = lane & 31;
  float wa[kSlots];
  {
    const float* __restrict__ wp = wslot + (size_t)p * kSlots * 32 + j;
#pragma unroll
    for (int r = 0; r < kSlots; ++r) wa[r] = wp[r * 32];
  }
  // output neuron of each of this lane's 16 D rows (-1 past the panel's rows)
  int orow[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = 8 * (q >> 2) + 4 * half + (q & 3);
    orow[q] = i < G ? outs[beg + i] : -1;
  }
  for (int64_t base = first; base < count; base += (int64_t)gridDim.y * kBlock) {
    // columns past the count read column 0 (in bounds: count > 0 here), zeroed after
    const int64_t pos = base + lane;
    const bool valid = pos < count;
    const int64_t src = valid ? (map ? (int64_t)map[pos] : pos) : 0;
    float x[kSlots];
#pragma unroll
    for (int r = 0; r < kSlots; ++r) x[r] = yin[(int64_t)pc[r] * ldin + src];
    bool nz = false;
#pragma unroll
    for (int r = 0; r < kSlots; ++r) {
      x[r] = valid ? x[r] : 0.f;
      nz |= x[r] != 0.f;
    }
    floatx32 acc;
#pragma unroll
    for (int q = 0; q < 32; ++q) acc[q] = 0.f;
    if (__ballot(nz) != 0ull) {
#pragma unroll
      for (int r = 0; r < kSlots; ++r)
        acc = __builtin_amdgcn_mfma_f32_32x32x1f32(wa[r], x[r], acc, 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int64_t pb = base + 32 * b + j;
      const bool vb = pb < count;
      int32_t mi = 0;
      if (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) mi = mcent[vb ? pb : 0];
      bool flag = false;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int o = orow[q];
        const bool ok = vb && o >= 0;
        const int oc = o >= 0 ? o : 0;
        const float bb = bias_vec ? bias_vec[oc] : bias_s;
        bool f = false;
        const float y = epilogue<EPI>(acc[16 * b + q], bb, clip, ok, oc, pb, mi, zc, kcap,
                                      eps_post, f);
        flag |= ok && f;
        if (ok) yout[(int64_t)oc * ldout + pb] = y;
      }
      // the column's rows are split over lanes j and j + 32
      const uint64_t m = __ballot(flag);
      const bool any = ((m >> j) & 1ull) | ((m >> (j + 32)) & 1ull);
      if (flags && half == 0 && vb && any) flags[pb] = 1;
    }
  }
}

// ---- wide ELL SpMM -------------------------------------------------------------------
// Rows with more than 32 nonzeros (unstructured nets, src/beyond/general.hpp:36-72 and
// src/beyond/SNICIT/kernel.hpp:103-139), and dense layers (cols == null: entry r reads input
// r; src/beyond/general.hpp:5-33,74-102).  Block (32-row chunk, column tile); each
// wavefront owns 64 columns and walks its rows with the fixed ascending-input chain, entries
// fetched 16 at a time with scalar loads.  Padded entries carry value 0 (a zero product
// leaves the chain unchanged).  Columns <= narrow: one thread per (row, column) pair.
template <int EPI>
__global__ __launch_bounds__(kBlock) void ell_spmm_kernel(
    const int32_t* __restrict__ cols, const float* __restrict__ vals, int32_t K, int32_t rows,
    const float* __restrict__ bias_vec, float bias_s, float clip, const float* __restrict__ yin,
    int64_t ldin, float* __restrict__ yout, int64_t ldout, const int32_t* __restrict__ count_ptr,
    const int32_t* __restrict__ map, int32_t* __restrict__ flags, float* __restrict__ zc,
    int32_t kcap, const int32_t* __restrict__ mcent, float eps_post, int32_t narrow) {
  const int r0 = blockIdx.x * 32, r1 = min(rows, r0 + 32);
  const int64_t count = *count_ptr;
  if (count <= narrow) {
    if (blockIdx.y != 0) return;
    const int G = r1 - r0;
    const int64_t pairs = (int64_t)G * count;
    for (int64_t q = threadIdx.x; q < pairs; q += kBlock) {
      const int o = r0 + (int)(q % G);
      const int64_t pos = q / G;
      const int64_t src = map ? (int64_t)map[pos] : pos;
      const float* __restrict__ v = vals + (size_t)o * K;
      const int32_t* __restrict__ c = cols ? cols + (size_t)o * K : nullptr;
      float acc = 0.f;
      for (int r = 0; r < K; ++r)
        acc = __builtin_fmaf(v[r], yin[(int64_t)(c ? c[r] : r) * ldin + src], acc);
      const float b = bias_vec ? bias_vec[o] : bias_s;
      const int32_t mi = (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) ? mcent[pos] : 0;
      bool flag = false;
      const float y = epilogue<EPI>(acc, b, clip, true, o, pos, mi, zc, kcap, eps_post, flag);
      yout[(int64_t)o * ldout + pos] = y;
      if (flags && flag) flags[pos] = 1;
    }
    return;
  }
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  int64_t run0, run1;
  tile_run(count, run0, run1);
  for (int64_t base = run0 + wv * kWave; base < run1; base += kBlock) {
    const int64_t pos = base + lane;
    const bool valid = pos < count;
    const int64_t src = valid ? (map ? (int64_t)map[pos] : pos) : 0;
    int32_t mi = 0;
    if (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) mi = mcent[valid ? pos : 0];
    bool flag = false;
    for (int o = r0; o < r1; ++o) {
      const float* __restrict__ v = vals + (size_t)o * K;
      const int32_t* __restrict__ c = cols ? cols + (size_t)o * K : nullptr;
      float acc = 0.f;
      for (int r = 0; r < K; r += 16) {
        float x[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = yin[(int64_t)(c ? c[r + q] : r + q) * ldin + src];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc = __builtin_fmaf(v[r + q], valid ? x[q] : 0.f, acc);
      }
      const float b = bias_vec ? bias_vec[o] : bias_s;
      const float y = epilogue<EPI>(acc, b, clip, valid, o, pos, mi, zc, kcap, eps_post, flag);
      if (valid) yout[(int64_t)o * ldout + pos] = y;
    }
    if (flags && valid && flag) flags[pos] = 1;
  }
}

// LDS-staged wide-row SpMM.  A block stages the input tile x_s[r][c] = Y[r][column c0+c] for
// all in_rows input neurons and TC columns in LDS (row stride TC+4 floats, so the float4
// reads of different random rows spread over the banks).  Row-parallel form (RPT): thread t
// owns output row blockIdx.y*256 + t and walks its entries in ascending input order, 4 at a
// time (int4/float4 loads of its own row, the next four prefetched while the current four
// are used), reading the TC inputs of each entry from LDS into TC running accumulators, so
// one weight load feeds TC multiply-adds and no input is gathered from L2 per entry.
// Few-rows form (!RPT, layers with fewer rows than a block has threads): thread (c, j) owns
// column c of rows j, j + 256/TC, ...  Same chain and epilogues as ell_spmm_kernel.
// Per-column flags are ORed over the block's rows with ballots.
template <int TC>
__device__ __forceinline__ void ell_stage_tile(const float* __restrict__ yin, int64_t ldin,
                                               int32_t in_rows, int64_t c0, int64_t count,
                                               const int32_t* __restrict__ map,
                                               float* __restrict__ x_s) {
  constexpr int XS = TC + 4;
  for (int idx = threadIdx.x; idx < in_rows * TC; idx += 256) {
    const int r = idx / TC, c = idx % TC;
    const int64_t pos = c0 + c;
    float v = 0.f;
    if (pos < count) v = yin[(int64_t)r * ldin + (map ? (int64_t)map[pos] : pos)];
    x_s[r * XS + c] = v;
  }
}

template <int EPI, int TC, bool RPT>
__global__ __launch_bounds__(256) void ell_lds_kernel(
    const int32_t* __restrict__ cols, const float* __restrict__ vals, int32_t K, int32_t rows,
    int32_t in_rows, const float* __restrict__ bias_vec, float bias_s, float clip,
    const float* __restrict__ yin, int64_t ldin, float* __restrict__ yout, int64_t ldout,
    const int32_t* __restrict__ count_ptr, const int32_t* __restrict__ map,
    int32_t* __restrict__ flags, float* __restrict__ zc, int32_t kcap,
    const int32_t* __restrict__ mcent, float eps_post) {
  constexpr int XS = TC + 4;
  extern __shared__ float x_s[];  // [in_rows][XS]
  const int64_t count = *count_ptr;
  const int lane = lane_id();
  for (int64_t c0 = (int64_t)blockIdx.x * TC; c0 < count; c0 += (int64_t)gridDim.x * TC) {
    __syncthreads();  // the previous tile's reads are done
    ell_stage_tile<TC>(yin, ldin, in_rows, c0, count, map, x_s);
    __syncthreads();
    if (RPT) {
      const int o = blockIdx.y * 256 + threadIdx.x;
      uint32_t fbits = 0;
      if (o < rows) {
        float acc[TC];
#pragma unroll
        for (int c = 0; c < TC; ++c) acc[c] = 0.f;
        const float* __restrict__ vr = vals + (size_t)o * K;
        const int32_t* __restrict__ cr = cols ? cols + (size_t)o * K : nullptr;
        float4 v4 = *reinterpret_cast<const float4*>(vr);
        int4 c4 = cr ? *reinterpret_cast<const int4*>(cr) : make_int4(0, 1, 2, 3);
        for (int r = 0; r < K; r += 4) {
          const float wv[4] = {v4.x, v4.y, v4.z, v4.w};
          const int cv[4] = {c4.x, c4.y, c4.z, c4.w};
          if (r + 4 < K) {
            v4 = *reinterpret_cast<const float4*>(vr + r + 4);
            c4 = cr ? *reinterpret_cast<const int4*>(cr + r + 4)
                    : make_int4(r + 4, r + 5, r + 6, r + 7);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float* __restrict__ xp = x_s + cv[q] * XS;
#pragma unroll
            for (int c = 0; c < TC; c += 4) {
              const float4 xv = *reinterpret_cast<const float4*>(xp + c);
              acc[c + 0] = __builtin_fmaf(wv[q], xv.x, acc[c + 0]);
              acc[c + 1] = __builtin_fmaf(wv[q], xv.y, acc[c + 1]);
              acc[c + 2] = __builtin_fmaf(wv[q], xv.z, acc[c + 2]);
              acc[c + 3] = __builtin_fmaf(wv[q], xv.w, acc[c + 3]);
            }
          }
        }
        const float b = bias_vec ? bias_vec[o] : bias_s;
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          const int64_t pos = c0 + c;
          const bool valid = pos < count;
          int32_t mi = 0;
          if (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) mi = mcent[valid ? pos : c0];
          bool f = false;
          const float y = epilogue<EPI>(acc[c], b, clip, valid, o, pos, mi, zc, kcap, eps_post, f);
          if (valid) yout[(int64_t)o * ldout + pos] = y;
          if (valid && f) fbits |= 1u << c;
        }
      }
      if (flags) {
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          const uint64_t m = __ballot((fbits >> c) & 1u);
          if (m != 0ull && lane == 0) flags[c0 + c] = 1;
        }
      }
    } else {
      const int c = threadIdx.x % TC;
      const int64_t pos = c0 + c;
      const bool valid = pos < count;
      int32_t mi = 0;
      if (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) mi = mcent[valid ? pos : c0];
      bool fl = false;
      for (int o = threadIdx.x / TC; o < rows; o += 256 / TC) {
        const float* __restrict__ vr = vals + (size_t)o * K;
        const int32_t* __restrict__ cr = cols ? cols + (size_t)o * K : nullptr;
        float acc = 0.f;
        for (int r = 0; r < K; r += 4) {
          const float4 v4 = *reinterpret_cast<const float4*>(vr + r);
          const int4 c4 = cr ? *reinterpret_cast<const int4*>(cr + r) : make_int4(r, r + 1, r + 2, r + 3);
          acc = __builtin_fmaf(v4.x, x_s[c4.x * XS + c], acc);
          acc = __builtin_fmaf(v4.y, x_s[c4.y * XS + c], acc);
          acc = __builtin_fmaf(v4.z, x_s[c4.z * XS + c], acc);
          acc = __builtin_fmaf(v4.w, x_s[c4.w * XS + c], acc);
        }
        const float b = bias_vec ? bias_vec[o] : bias_s;
        bool f = false;
        const float y = epilogue<EPI>(acc, b, clip, valid, o, pos, mi, zc, kcap, eps_post, f);
        if (valid) yout[(int64_t)o * ldout + pos] = y;
        fl |= valid && f;
      }
      if (flags && fl) flags[pos] = 1;
    }
  }
}

// (row, column) chains for launches with few live columns: thread `first` of `nthreads`
// takes pairs first, first + nthreads, ..., the lanes of a wavefront along the columns of
// one row and on to the next row, reading inputs and weights from L2.  Same ascending chain
// and epilogue per (row, column) as the tiled forms.
template <int EPI>
__device__ __forceinline__ void ell_pair_chains(
    const int32_t* __restrict__ cols, const float* __restrict__ vals, int32_t K, int32_t rows,
    const float* __restrict__ bias_vec, float bias_s, float clip, const float* __restrict__ yin,
    int64_t ldin, float* __restrict__ yout, int64_t ldout, int64_t count,
    const int32_t* __restrict__ map, int32_t* __restrict__ flags, float* __restrict__ zc,
    int32_t kcap, const int32_t* __restrict__ mcent, float eps_post, int64_t first,
    int64_t nthreads) {
  const int64_t pairs = (int64_t)rows * count;
  for (int64_t q = first; q < pairs; q += nthreads) {
    const int oo = (int)(q / count);
    const int64_t pos = q % count;
    const int64_t src = map ? (int64_t)map[pos] : pos;
    const float* __restrict__ v = vals + (size_t)oo * K;
    const int32_t* __restrict__ c = cols ? cols + (size_t)oo * K : nullptr;
    float acc = 0.f;
    int r = 0;
    for (; r + 8 <= K; r += 8) {
      float x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = yin[(int64_t)(c ? c[r + u] : r + u) * ldin + src];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = __builtin_fmaf(v[r + u], x[u], acc);
    }
    for (; r < K; ++r) acc = __builtin_fmaf(v[r], yin[(int64_t)(c ? c[r] : r) * ldin + src], acc);
    const float b = bias_vec ? bias_vec[oo] : bias_s;
    const int32_t mi = (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) ? mcent[pos] : 0;
    bool flag = false;
    const float y = epilogue<EPI>(acc, b, clip, true, oo, pos, mi, zc, kcap, eps_post, flag);
    yout[(int64_t)oo * ldout + pos] = y;
    if (flags && flag) flags[pos] = 1;
  }
}

// Chunked row-parallel form for wide tiles: thread t owns output row blockIdx.x*256 + t and
// TC = 64 columns.  The input rows are staged in chunks of rc rows (x_s[r - R0][c], row stride
// TC+4); a thread's chain pauses at the first entry past the chunk and resumes with the next
// chunk, so the multiply-add order over its ascending entries is unchanged.  Entries are read
// four at a time (int4/float4) and each feeds TC multiply-adds from LDS.  Consecutive blocks
// take the row groups of one column tile, so the tile's inputs come from L2 after the first.
template <int EPI>
__global__ __launch_bounds__(256) void ell_chunk_kernel(
    const int32_t* __restrict__ cols, const float* __restrict__ vals, int32_t K, int32_t rows,
    int32_t in_rows, int32_t rc, const float* __restrict__ bias_vec, float bias_s, float clip,
    const float* __restrict__ yin, int64_t ldin, float* __restrict__ yout, int64_t ldout,
    const int32_t* __restrict__ count_ptr, const int32_t* __restrict__ map,
    int32_t* __restrict__ flags, float* __restrict__ zc, int32_t kcap,
    const int32_t* __restrict__ mcent, float eps_post, int32_t narrow) {
  constexpr int TC = 64, XS = TC + 4;
  extern __shared__ float x_s[];  // [rc][XS]
  const int64_t count = *count_ptr;
  const int lane = lane_id();
  if (count <= narrow) {
    // Few columns (a small serving group): the row-per-thread form would keep only the
    // first column tile's workgroups busy, each running 64 column chains for a handful of
    // live columns.  Instead every workgroup of the grid takes (row, column) chains.
    ell_pair_chains<EPI>(cols, vals, K, rows, bias_vec, bias_s, clip, yin, ldin, yout, ldout,
                         count, map, flags, zc, kcap, mcent, eps_post,
                         ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 256 + threadIdx.x,
                         (int64_t)gridDim.x * gridDim.y * 256);
    return;
  }
  const int o = blockIdx.x * 256 + threadIdx.x;
  const bool row_ok = o < rows;
  const float* __restrict__ vr = vals + (size_t)(row_ok ? o : 0) * K;
  const int32_t* __restrict__ cr = cols ? cols + (size_t)(row_ok ? o : 0) * K : nullptr;
  for (int64_t c0 = (int64_t)blockIdx.y * TC; c0 < count; c0 += (int64_t)gridDim.y * TC) {
    float acc[TC];
#pragma unroll
    for (int c = 0; c < TC; ++c) acc[c] = 0.f;
    int p = 0;   // next group of four entries
    int bi = 4;  // position in the loaded group (4: none loaded)
    int4 cb = make_int4(0, 0, 0, 0);
    float4 vb = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int R0 = 0; R0 < in_rows; R0 += rc) {
      const int nr = min(rc, in_rows - R0);
      __syncthreads();  // the previous chunk's reads are done
      for (int idx = threadIdx.x; idx < nr * TC; idx += 256) {
        const int r = idx / TC, c = idx % TC;
        const int64_t pos = c0 + c;
        float v = 0.f;
        if (pos < count) v = yin[(int64_t)(R0 + r) * ldin + (map ? (int64_t)map[pos] : pos)];
        x_s[r * XS + c] = v;
      }
      __syncthreads();
      if (!row_ok) continue;
      const int R1 = R0 + nr;
      while (true) {
        if (bi == 4) {
          if (p >= K) break;
          vb = *reinterpret_cast<const float4*>(vr + p);
          cb = cr ? *reinterpret_cast<const int4*>(cr + p) : make_int4(p, p + 1, p + 2, p + 3);
          bi = 0;
          p += 4;
        }
        const int col = bi == 0 ? cb.x : bi == 1 ? cb.y : bi == 2 ? cb.z : cb.w;
        if (col >= R1) break;
        const float w = bi == 0 ? vb.x : bi == 1 ? vb.y : bi == 2 ? vb.z : vb.w;
        const float* __restrict__ xp = x_s + (col - R0) * XS;
#pragma unroll
        for (int c = 0; c < TC; c += 4) {
          const float4 xv = *reinterpret_cast<const float4*>(xp + c);
          acc[c + 0] = __builtin_fmaf(w, xv.x, acc[c + 0]);
          acc[c + 1] = __builtin_fmaf(w, xv.y, acc[c + 1]);
          acc[c + 2] = __builtin_fmaf(w, xv.z, acc[c + 2]);
          acc[c + 3] = __builtin_fmaf(w, xv.w, acc[c + 3]);
        }
        ++bi;
      }
    }
    uint64_t fbits = 0;
    if (row_ok) {
      const float b = bias_vec ? bias_vec[o] : bias_s;
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        const int64_t pos = c0 + c;
        const bool valid = pos < count;
        int32_t mi = 0;
        if (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) mi = mcent[valid ? pos : c0];
        bool f = false;
        const float y = epilogue<EPI>(acc[c], b, clip, valid, o, pos, mi, zc, kcap, eps_post, f);
        if (valid) yout[(int64_t)o * ldout + pos] = y;
        if (valid && f) fbits |= 1ull << c;
      }
    }
    if (flags) {
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        const uint64_t m = __ballot((fbits >> c) & 1ull);
        if (m != 0ull && lane == 0) flags[c0 + c] = 1;
      }
    }
  }
}

// Wavefront-per-row wide-row SpMM.  A work item is (column tile, block of rows): the tile is
// 64*W columns, lane l owning columns W*l .. W*l+W-1, so every LDS read of an input neuron
// is one conflict-free run (W floats per lane, one ds_read_b32 or ds_read_b128).  Each entry
// is wave-uniform: an (input neuron - first neuron of its chunk, weight) pair of the layer's
// padded stream, read kEllGroup at a time and used for W columns per lane, so the cost of
// broadcasting it is shared by 64*W multiply-adds.  The tile's inputs are staged in chunks
// of kEllChunk neurons, two chunks resident: while the block runs chunk c from one buffer,
// chunk c+1 is already loaded into registers and is written to the other buffer afterwards,
// one barrier per chunk.  A row's entries that read chunk c are [rng[o][c], rng[o][c+1])
// of the stream (ascending input, padded to whole groups with value-0 entries), so the chain
// of every column pauses at each chunk boundary and resumes with the next chunk: the
// multiply-add order is that of the oracle.  Rows per wavefront (rw <= kEllRW) come from the
// live column count, so small batches still spread over every CU.  With at least 8 column
// tiles, the row blocks of one tile are work items of one XCD (workgroup i runs on XCD
// i mod 8), so the tile is fetched into that XCD's L2 once.
constexpr int kEllWaves = 16;
constexpr int kEllRW = 8;
constexpr int kEllStageRows = kEllChunk / kEllWaves;  // input rows each wavefront stages

template <int W> struct ColVec { typedef float4 type; };
template <> struct ColVec<1> { typedef float type; };
__device__ __forceinline__ float col_of(float v, int) { return v; }
__device__ __forceinline__ float col_of(const float4& v, int j) {
  return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
}

template <int EPI, int W>
__global__ __launch_bounds__(kEllWaves * 64) void ell_wave_kernel(
    const int2* __restrict__ ent, int32_t rows, int32_t in_rows,
    const int32_t* __restrict__ rng, const float* __restrict__ bias_vec, float bias_s,
    float clip, const float* __restrict__ yin, int64_t ldin, float* __restrict__ yout,
    int64_t ldout, const int32_t* __restrict__ count_ptr, const int32_t* __restrict__ map,
    int32_t* __restrict__ flags, float* __restrict__ zc, int32_t kcap,
    const int32_t* __restrict__ mcent, float eps_post, int32_t min_items) {
  static_assert(W == 1 || W == 4, "1 or 4 columns per lane");
  typedef typename ColVec<W>::type Vec;
  constexpr int TC = 64 * W;  // columns per tile
  extern __shared__ float xs[];  // [2][kEllChunk][TC]
  const int64_t count = *count_ptr;
  const int lane = lane_id();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int nch = (in_rows + kEllChunk - 1) / kEllChunk;
  const int64_t tiles = (count + TC - 1) / TC;
  // rows per wavefront: the largest power of two <= kEllRW that still gives min_items
  int rw = kEllRW;
  while (rw > 1 && tiles * ((rows + kEllWaves * rw - 1) / (kEllWaves * rw)) < min_items) rw >>= 1;
  const int64_t rbc = (rows + kEllWaves * rw - 1) / (kEllWaves * rw);
  const bool xcd_major = tiles >= 8;
  const int64_t items = (xcd_major ? (tiles + 7) / 8 * 8 : tiles) * rbc;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    int64_t tile, rb;
    if (xcd_major) {
      const int64_t jj = it / 8;
      tile = jj / rbc * 8 + it % 8;
      rb = jj % rbc;
    } else {
      tile = it / rbc;
      rb = it % rbc;
    }
    if (tile >= tiles) continue;  // uniform over the block
    int64_t pos[W], src[W];
    bool valid[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
      pos[j] = tile * TC + lane * W + j;
      valid[j] = pos[j] < count;
      src[j] = valid[j] ? (map ? (int64_t)map[pos[j]] : pos[j]) : 0;
    }
    const int row0 = (int)(rb * kEllWaves * rw) + wv * rw;
    float st[kEllStageRows][W];
    auto load_chunk = [&](int c) {
#pragma unroll
      for (int q = 0; q < kEllStageRows; ++q) {
        const int r = c * kEllChunk + q * kEllWaves + wv;
#pragma unroll
        for (int j = 0; j < W; ++j)
          st[q][j] = (valid[j] && r < in_rows) ? yin[(int64_t)r * ldin + src[j]] : 0.f;
      }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
      for (int q = 0; q < kEllStageRows; ++q) {
        float* __restrict__ x = xs + buf * (kEllChunk * TC) + (q * kEllWaves + wv) * TC + lane * W;
        if constexpr (W == 4) *reinterpret_cast<float4*>(x) = make_float4(st[q][0], st[q][1], st[q][2], st[q][3]);
        else x[0] = st[q][0];
      }
    };
    float acc[kEllRW][W];
#pragma unroll
    for (int i = 0; i < kEllRW; ++i)
#pragma unroll
      for (int j = 0; j < W; ++j) acc[i][j] = 0.f;
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
      if (c + 1 < nch) load_chunk(c + 1);
      // this lane's columns in the chunk's buffer; an entry adds its input row's offset
      const char* __restrict__ xb =
          reinterpret_cast<const char*>(xs + (c & 1) * (kEllChunk * TC) + lane * W);
      // every row's entry range in this chunk, loaded together rather than one row at a time
      // just before its entries (wide-row BF layers of the 784->1024x24 net: 2.83 against
      // 2.98 ms per layer at ~132k columns)
      int pr[kEllRW], er[kEllRW];
#pragma unroll
      for (int i = 0; i < kEllRW; ++i) {
        const int o = row0 + i;
        const bool in = i < rw && o < rows;
        const int32_t* __restrict__ rr = rng + (size_t)(in ? o : 0) * (nch + 1);
        pr[i] = in ? rr[c] : 0;
        er[i] = in ? rr[c + 1] : 0;
      }
      int2 pre[kEllGroup];
      bool have = false;  // pre holds row i's first group (loaded by row i - 1's last step)
#pragma unroll
      for (int i = 0; i < kEllRW; ++i) {
        {
          int p = pr[i];
          const int e = er[i];
          // the step after a row's last group loads the next row's first group instead of
          // slack (the last row's runs on into the stream's slack, as before): 64.1 -> 60.6
          // ms for the plain batch of the 784->1024x24 net
          const int pnext = i + 1 < kEllRW ? pr[i + 1] : e;
          if (p >= e) have = false;
          if (p < e) {
            // whole groups (runs are padded), two register sets in turn: the next group's
            // load goes out before the current group's LDS reads (the stream carries a
            // group of slack past its end)
            float a[W];
#pragma unroll
            for (int j = 0; j < W; ++j) a[j] = acc[i][j];
            int2 ga[kEllGroup], gb[kEllGroup];
            auto load = [&](int2* g, int at) {
#pragma unroll
              for (int q = 0; q < kEllGroup; ++q) g[q] = ent[at + q];
            };
            auto group = [&](const int2* g) {
              Vec xv[kEllGroup];
#pragma unroll
              for (int q = 0; q < kEllGroup; ++q)
                xv[q] = *reinterpret_cast<const Vec*>(xb + g[q].x * (TC * (int)sizeof(float)));
#pragma unroll
              for (int q = 0; q < kEllGroup; ++q)
#pragma unroll
                for (int j = 0; j < W; ++j)
                  a[j] = __builtin_fmaf(__int_as_float(g[q].y), col_of(xv[q], j), a[j]);
            };
            if (have) {
#pragma unroll
              for (int q = 0; q < kEllGroup; ++q) ga[q] = pre[q];
            } else {
              load(ga, p);
            }
            while (true) {
              p += kEllGroup;
              load(gb, p < e ? p : pnext);
              __builtin_amdgcn_sched_barrier(0);
              group(ga);
              if (p >= e) {
#pragma unroll
                for (int q = 0; q < kEllGroup; ++q) pre[q] = gb[q];
                break;
              }
              p += kEllGroup;
              load(ga, p < e ? p : pnext);
              __builtin_amdgcn_sched_barrier(0);
              group(gb);
              if (p >= e) {
#pragma unroll
                for (int q = 0; q < kEllGroup; ++q) pre[q] = ga[q];
                break;
              }
            }
            have = true;
#pragma unroll
            for (int j = 0; j < W; ++j) acc[i][j] = a[j];
          }
        }
      }
      if (c + 1 < nch) store_chunk((c + 1) & 1);
      __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < W; ++j) {
      int32_t mi = 0;
      if (EPI == EPI_DELTA || EPI == EPI_DELTA_LOSSY) mi = valid[j] ? mcent[pos[j]] : 0;
      bool flag = false;
#pragma unroll
      for (int i = 0; i < kEllRW; ++i) {
        const int o = row0 + i;
        if (i < rw && o < rows) {
          const float b = bias_vec ? bias_vec[o] : bias_s;
          const float y =
              epilogue<EPI>(acc[i][j], b, clip, valid[j], o, pos[j], mi, zc, kcap, eps_post, flag);
          if (valid[j]) yout[(int64_t)o * ldout + pos[j]] = y;
        }
      }
      if (flags && valid[j] && flag) flags[pos[j]] = 1;
    }
  }
}

__global__ __launch_bounds__(256) void argmax_kernel(const float* __restrict__ logits, int64_t ld,
                                                      int32_t rows,
                                                      const int32_t* __restrict__ count_ptr,
                                                      int32_t* __restrict__ cat) {
  const int64_t count = *count_ptr;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  // check_acc, src/beyond/general.hpp:112-119: first maximum, starting from -10000
  float best = -10000.f;
  int arg = 0;
  for (int o = 0; o < rows; ++o) {
    const float v = logits[(int64_t)o * ld + j];
    if (v > best) {
      best = v;
      arg = o;
    }
  }
  cat[j] = arg;
}

// ---- batch segments ---------------------------------------------------------------------
// A plan's columns can hold several independent batches ("segments"), each clustered on its
// own.  segstart[s] is the first original column of segment s; the live list keeps original
// columns ascending, so segment s occupies list positions [seglo[s], seglo[s+1]).  A plain
// plan is the one-segment case (segstart = {0, INT_MAX}).
__global__ void segment_bounds_kernel(const int32_t* __restrict__ orig,
                                      const int32_t* __restrict__ count_ptr,
                                      const int32_t* __restrict__ segstart, int32_t nseg,
                                      int32_t* __restrict__ seglo) {
  const int count = *count_ptr;
  for (int s = threadIdx.x; s <= nseg; s += blockDim.x) {
    if (s == nseg) {
      seglo[s] = count;
      continue;
    }
    const int32_t key = segstart[s];
    int lo = 0, hi = count;  // first position with orig >= key
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (orig[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    seglo[s] = lo;
  }
}

// Segment of list position i: the last s with seglo[s] <= i (empty segments skipped).
__device__ __forceinline__ int segment_of(const int32_t* __restrict__ seglo, int nseg,
                                          int64_t i) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seglo[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---- lossy clustering ------------------------------------------------------------------
// Pool (y_star_gen, src/beyond/SNICIT/kernel.hpp:12): pool row y of a batch segment with
// B_s rows is its row y*B_s/pool.  Every row is live in this variant, so list positions are
// columns.  Block (pool row, batch segment).
__global__ __launch_bounds__(256) void lossy_pool_kernel(const float* __restrict__ Y, int64_t ld,
                                                          const int32_t* __restrict__ seglo,
                                                          int32_t pool, int32_t N,
                                                          float* __restrict__ poolrows,
                                                          int32_t* __restrict__ poolcol) {
  const int y = blockIdx.x, sg = blockIdx.y;
  const int64_t lo = seglo[sg], Bs = seglo[sg + 1] - lo;
  const int64_t col = Bs > 0 ? lo + (int64_t)y * Bs / pool : -1;
  const int64_t row = (int64_t)sg * pool + y;
  if (threadIdx.x == 0) poolcol[row] = (int32_t)col;
  for (int n = threadIdx.x; n < N; n += blockDim.x)
    poolrows[row * N + n] = col >= 0 ? Y[(int64_t)n * ld + col] : 0.f;
}

// Pruning (y_star_gen, src/beyond/SNICIT/kernel.hpp:19-41): in pool order, every present
// base i removes each other present row y with #{n : |P_y[n] - P_i[n]| > eps} < N*eps.
// Two kernels.  lossy_close_kernel: workgroup (row i, batch segment) decides for every later
// pool row y > i whether it is that close to row i, wavefront w testing rows i+1+w,
// i+17+w, ...; lanes stride the neurons, 256 per step, and a pair stops as soon as its count
// reaches N*eps (the count only grows), so most pairs of different rows end after a step or
// two.  The pair tests are independent, so every pool row's tests run at once.  Row i's
// 128-bit mask of close later rows goes to close[segment][i].  lossy_prune_kernel: one
// wavefront per segment walks the rows in pool order; a row still present removes the rows
// in its mask.  A present earlier row close to row i would have removed i, so later rows are
// the only ones the walk can remove, and the survivors are those of the sequential rule.
__global__ __launch_bounds__(1024) void lossy_close_kernel(const float* __restrict__ poolrows,
                                                          const int32_t* __restrict__ seglo,
                                                          int32_t pool, int32_t N, float cut,
                                                          double limit,
                                                          uint64_t* __restrict__ close) {
  __shared__ uint64_t bits[2];
  const int i = blockIdx.x, sg = blockIdx.y;
  const int t = threadIdx.x, lane = lane_id(), wv = t / kWave;
  const int64_t Bs = seglo[sg + 1] - seglo[sg];
  const int P = Bs > 0 ? pool : 0;
  if (t < 2) bits[t] = 0ull;
  __syncthreads();
  if (i < P) {
    const float* __restrict__ rows = poolrows + (int64_t)sg * pool * N;
    const float* __restrict__ pi = rows + (int64_t)i * N;
    for (int y = i + 1 + wv; y < P; y += 16) {
      const float* __restrict__ py = rows + (int64_t)y * N;
      int cnt = 0;
      for (int n0 = 0; n0 < N && (double)cnt < limit; n0 += 4 * kWave) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int n = n0 + u * kWave + lane;
          const bool d = n < N && fabsf(py[n] - pi[n]) > cut;
          cnt += __popcll(__ballot(d));
        }
      }
      if ((double)cnt < limit && lane == 0) atomicOr(&bits[y >> 6], 1ull << (y & 63));
    }
  }
  __syncthreads();
  if (t < 2) close[((int64_t)sg * kMaxCentroidsLossy + i) * 2 + t] = bits[t];
}

__global__ __launch_bounds__(kWave) void lossy_prune_kernel(
    const uint64_t* __restrict__ close, const int32_t* __restrict__ poolcol,
    const int32_t* __restrict__ seglo, int32_t pool, int32_t stride,
    int32_t* __restrict__ cent_col, int32_t* __restrict__ kseg, int32_t* __restrict__ kcols) {
  __shared__ uint64_t m[kMaxCentroidsLossy * 2];
  const int sg = blockIdx.x, lane = threadIdx.x;
  const int64_t Bs = seglo[sg + 1] - seglo[sg];
  const int P = Bs > 0 ? pool : 0;
  const uint64_t* __restrict__ cs = close + (int64_t)sg * kMaxCentroidsLossy * 2;
  for (int q = lane; q < 2 * P; q += kWave) m[q] = cs[q];
  __syncthreads();
  // present rows: bits 0..P-1 of (a0, a1)
  uint64_t a0 = P >= 64 ? ~0ull : (1ull << P) - 1ull;
  uint64_t a1 = P >= 128 ? ~0ull : P > 64 ? (1ull << (P - 64)) - 1ull : 0ull;
  for (int i = 0; i < P; ++i) {
    const bool present = ((i < 64 ? a0 >> i : a1 >> (i - 64)) & 1ull) != 0ull;
    if (present) {
      a0 &= ~m[2 * i];
      a1 &= ~m[2 * i + 1];
    }
  }
  if (lane == 0) {
    int k = 0;
    for (int y = 0; y < P; ++y)
      if ((y < 64 ? a0 >> y : a1 >> (y - 64)) & 1ull)
        cent_col[sg * stride + k++] = poolcol[sg * pool + y];
    kseg[sg] = k;
    if (k > 0) atomicMax(kcols, sg * stride + k);
  }
}

// Assignment (coarse_cluster, src/beyond/SNICIT/kernel.hpp:63-87): distance = #{n :
// |C_c[n] - Y_i[n]| > eps} over the centroids of i's batch segment, first minimum; centroid
// columns map to their own index.  Workgroup (64 columns, block of eight centroids): lane =
// column, wavefront w counts neurons [w N/16, (w+1) N/16) with eight neurons' loads issued
// before their compares; the 16 partial counts of a (column, centroid) are summed in LDS and
// stored once, centroid-major (dist[c][i], row stride ld; integers: the order does not
// matter).  Lanes of different batch segments are
// handled one segment at a time, as in the exact variant's walk.
constexpr int kDistWaves = 16;

__global__ __launch_bounds__(kDistWaves * kWave) void lossy_dist_kernel(
    const float* __restrict__ Y, int64_t ld, const int32_t* __restrict__ seglo, int32_t nseg,
    const float* __restrict__ C, int32_t kcap, int32_t stride, const int32_t* __restrict__ kseg,
    int32_t N, float eps, int32_t* __restrict__ dist) {
  __shared__ int32_t part[kDistWaves][8][kWave];
  const int64_t count = seglo[nseg];
  if ((int64_t)blockIdx.x * kWave >= count) return;  // the whole workgroup
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  const int64_t i = (int64_t)blockIdx.x * kWave + lane;  // < ld: ld is a multiple of 64
  const bool valid = i < count;
  const int sg = !valid ? -1 : nseg == 1 ? 0 : segment_of(seglo, nseg, i);
  const int cb = 8 * (int)blockIdx.y;
  const int n0 = (int)((int64_t)N * wv / kDistWaves);
  const int n1 = (int)((int64_t)N * (wv + 1) / kDistWaves);
  int tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t todo = __ballot(valid);
  while (todo) {
    const int lead = __ffsll((unsigned long long)todo) - 1;
    const int s0 = __builtin_amdgcn_readfirstlane(__shfl(sg, lead));
    const bool mine = valid && sg == s0;
    todo &= ~__ballot(mine);
    if (cb >= kseg[s0]) continue;
    const float* __restrict__ Cs = C + (int64_t)s0 * stride + cb;
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int n = n0;
    for (; n + 8 <= n1; n += 8) {
      float y[8], c[8][8];
#pragma unroll
      for (int u = 0; u < 8; ++u) y[u] = Y[(int64_t)(n + u) * ld + i];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int q = 0; q < 8; ++q) c[u][q] = Cs[(int64_t)(n + u) * kcap + q];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int q = 0; q < 8; ++q) cnt[q] += fabsf(c[u][q] - y[u]) > eps;
    }
    for (; n < n1; ++n) {
      const float y = Y[(int64_t)n * ld + i];
#pragma unroll
      for (int q = 0; q < 8; ++q) cnt[q] += fabsf(Cs[(int64_t)n * kcap + q] - y) > eps;
    }
    if (mine) {
#pragma unroll
      for (int q = 0; q < 8; ++q) tot[q] += cnt[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) part[wv][q][lane] = tot[q];
  __syncthreads();
  if (wv == 0 && valid) {
    const int k = kseg[sg];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (cb + q >= k) break;
      int sum = 0;
#pragma unroll
      for (int w = 0; w < kDistWaves; ++w) sum += part[w][q][lane];
      dist[(int64_t)(cb + q) * ld + i] = sum;
    }
  }
}

__global__ __launch_bounds__(256) void lossy_final_kernel(
    const int32_t* __restrict__ seglo, int32_t nseg, const int32_t* __restrict__ kseg,
    int32_t stride, const int32_t* __restrict__ cent_col, const int32_t* __restrict__ dist,
    int64_t ld, int32_t* __restrict__ mcar, int32_t* __restrict__ centmark,
    int32_t* __restrict__ rec) {
  const int64_t count = seglo[nseg];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int sg = segment_of(seglo, nseg, i);
  const int k = kseg[sg];
  const int32_t* __restrict__ cs = cent_col + sg * stride;
  int best = 0x7fffffff, bestc = 0;
  // sixteen loads in flight, then the first minimum in centroid order
  for (int c0 = 0; c0 < k; c0 += 16) {
    int d[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) d[u] = c0 + u < k ? dist[(int64_t)(c0 + u) * ld + i] : 0x7fffffff;
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (d[u] < best) {
        best = d[u];
        bestc = c0 + u;
      }
  }
  int mark = 0;
  for (int c = 0; c < k; ++c)
    if (cs[c] == i && !mark) {
      mark = 1;
      bestc = c;
    }
  const int32_t m = sg * stride + bestc;
  mcar[i] = m;
  centmark[i] = mark;
  rec[i] = k > 0 ? m : -1;
}

// Delta (coarse_cluster, src/beyond/SNICIT/kernel.hpp:89-99): v = |Y_i - C_m| > eps ?
// Y_i - C_m : 0; non-empty iff any v != 0 (any v > 0 with positive_only, as the reference).
__global__ __launch_bounds__(256) void lossy_delta_kernel(
    const float* __restrict__ Y, int64_t ld, const float* __restrict__ C, int32_t kcap,
    const int32_t* __restrict__ mcar, const int32_t* __restrict__ centmark,
    const int32_t* __restrict__ B_ptr, int32_t N, float eps, int32_t positive_only,
    float* __restrict__ D, int64_t ldD, int32_t* __restrict__ flags) {
  const int64_t count = *B_ptr;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count || centmark[i]) return;
  const int c = mcar[i];
  const int n0 = blockIdx.y * 64, n1 = min(N, n0 + 64);
  bool any = false;
#pragma unroll 8
  for (int n = n0; n < n1; ++n) {
    const float d = Y[(int64_t)n * ld + i] - C[(int64_t)n * kcap + c];
    const float v = fabsf(d) > eps ? d : 0.f;
    D[(int64_t)n * ldD + i] = v;
    any |= positive_only ? (v > 0.f) : (v != 0.f);
  }
  if (any) flags[i] = 1;
}

// ---- stable compaction ----------------------------------------------------------------
// A workgroup of 1024 threads walks flags in rounds of SUB x 1024 columns: SUB coalesced
// loads per thread are in flight together, a ballot per 1024-column sub-round gives
// per-wavefront counts, one wavefront scans the 16 SUB counts, and every set flag gets its
// output slot, so the new list keeps column order.  Flags are cleared as they are consumed.
// compose (optional) maps each flagged column to the value stored in map_out; it must not
// alias map_out.  Up to kCompactOneBlock columns, one workgroup walks them all (16 x 1024
// per round).  Beyond that, compact_count_kernel counts the set flags of every 4096-column
// tile, and workgroup b of compact_kernel places tile b's columns after the sum of the
// earlier tiles' counts: the same order, with the tiles' rounds side by side and no waiting
// between workgroups.  Replaces the per-layer host loops over managed memory
// (src/SDGC/BF/BF.hpp:222-229, src/SDGC/SNICIT/SNICIT.hpp:556-615).
constexpr int kCompactSub = 16;
constexpr int kCompactTileSub = 4;
constexpr int kCompactTile = 1024 * kCompactTileSub;
constexpr int kCompactOneBlock = 1024 * kCompactSub;

// Rounds over columns [begin, end), output slots from `running`; returns the slot after the
// last one placed.  Every thread of the workgroup calls it with the same arguments.
template <int SUB>
__device__ __forceinline__ int compact_range(
    int32_t* __restrict__ flags, int begin, int end, int running, int32_t* __restrict__ map_out,
    const int32_t* __restrict__ compose, const int32_t* __restrict__ c0_in,
    int32_t* __restrict__ c0_out, const int32_t* __restrict__ c1_in,
    int32_t* __restrict__ c1_out, int32_t* s_cnt, int32_t* s_off, int32_t* s_tot) {
  constexpr int E = SUB * 16;   // (sub-round, wavefront) entries, in order e = 16 q + w
  constexpr int EPL = E / 64;   // entries per lane of the scanning wavefront
  const int t = threadIdx.x, lane = lane_id(), wv = t / kWave;
  for (int base = begin; base < end; base += 1024 * SUB) {
    // every load of the round is issued before the scan: flags and the carried values
    bool f[SUB];
    int32_t cm[SUB], v0[SUB], v1[SUB];
#pragma unroll
    for (int q = 0; q < SUB; ++q) {
      const int i = base + q * 1024 + t;
      const bool in = i < end;
      f[q] = in && flags[i] != 0;
      cm[q] = (in && compose) ? compose[i] : i;
      v0[q] = (in && c0_in) ? c0_in[i] : 0;
      v1[q] = (in && c1_in) ? c1_in[i] : 0;
    }
    uint64_t m[SUB];
#pragma unroll
    for (int q = 0; q < SUB; ++q) {
      m[q] = __ballot(f[q]);
      if (lane == 0) s_cnt[q * 16 + wv] = __popcll(m[q]);
    }
    __syncthreads();
    if (wv == 0) {
      int v[EPL], tot = 0;
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        v[e] = s_cnt[EPL * lane + e];
        tot += v[e];
      }
      int incl = tot;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int x = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += x;
      }
      int ex = incl - tot;
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        s_off[EPL * lane + e] = ex;
        ex += v[e];
      }
      if (lane == kWave - 1) *s_tot = incl;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SUB; ++q) {
      if (f[q]) {
        const int o = running + s_off[q * 16 + wv] + lane_prefix(m[q]);
        map_out[o] = cm[q];
        if (c0_out) c0_out[o] = v0[q];
        if (c1_out) c1_out[o] = v1[q];
        flags[base + q * 1024 + t] = 0;
      }
    }
    running += *s_tot;
    __syncthreads();
  }
  return running;
}

__global__ __launch_bounds__(1024) void compact_count_kernel(const int32_t* __restrict__ flags,
                                                            const int32_t* __restrict__ count_in,
                                                            int32_t* __restrict__ tile_cnt) {
  __shared__ int32_t s_sum;
  const int n = *count_in;
  const int64_t base = (int64_t)blockIdx.x * kCompactTile;
  if (n <= kCompactOneBlock || base >= n) return;
  const int t = threadIdx.x, lane = lane_id();
  if (t == 0) s_sum = 0;
  __syncthreads();
  int c = 0;
#pragma unroll
  for (int q = 0; q < kCompactTileSub; ++q) {
    const int64_t i = base + q * 1024 + t;
    c += __popcll(__ballot(i < n && flags[i] != 0));
  }
  if (lane == 0 && c) atomicAdd(&s_sum, c);
  __syncthreads();
  if (t == 0) tile_cnt[blockIdx.x] = s_sum;
}

// The list kept as it is (slack compaction): map_out[i] = compose[i] (or i) and the carried
// values copied for every position of [begin, end); flags cleared.
__device__ __forceinline__ void keep_range(int32_t* __restrict__ flags, int begin, int end,
                                           int32_t* __restrict__ map_out,
                                           const int32_t* __restrict__ compose,
                                           const int32_t* __restrict__ c0_in,
                                           int32_t* __restrict__ c0_out,
                                           const int32_t* __restrict__ c1_in,
                                           int32_t* __restrict__ c1_out) {
  for (int i = begin + (int)threadIdx.x; i < end; i += (int)blockDim.x) {
    map_out[i] = compose ? compose[i] : i;
    if (c0_out) c0_out[i] = c0_in ? c0_in[i] : 0;
    if (c1_out) c1_out[i] = c1_in ? c1_in[i] : 0;
    flags[i] = 0;
  }
}

// slack > 0: when at most n * slack / 1024 of the n listed columns are unflagged, the list is
// kept as it is (count_out = n) instead of compacted, so the next layer reads its columns in
// place, aligned; the unflagged columns ride along until a later compaction drops them.
// live_out (nullable) receives the flagged count either way.
__global__ __launch_bounds__(1024) void compact_kernel(
    int32_t* __restrict__ flags, const int32_t* __restrict__ count_in,
    int32_t* __restrict__ count_out, int32_t* __restrict__ map_out,
    const int32_t* __restrict__ compose, const int32_t* __restrict__ c0_in,
    int32_t* __restrict__ c0_out, const int32_t* __restrict__ c1_in,
    int32_t* __restrict__ c1_out, const int32_t* __restrict__ tile_cnt, int32_t slack,
    int32_t* __restrict__ live_out) {
  __shared__ int32_t s_cnt[kCompactSub * 16];
  __shared__ int32_t s_off[kCompactSub * 16];
  __shared__ int32_t s_tot;
  __shared__ int32_t s_base;
  __shared__ int32_t s_live;
  const int n = *count_in;
  const int t = threadIdx.x, lane = lane_id();
  const auto keep_list = [&](int live) {
    return slack > 0 && (int64_t)(n - live) * 1024 <= (int64_t)n * slack;
  };
  if (n <= kCompactOneBlock || gridDim.x == 1) {  // one workgroup walks the whole list
    if (blockIdx.x != 0) return;
    if (slack > 0) {
      if (t == 0) s_live = 0;
      __syncthreads();
      int c = 0;
      for (int i = t; i < n; i += 1024) c += flags[i] != 0;
      for (int d = kWave / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, kWave);
      if (lane == 0 && c) atomicAdd(&s_live, c);
      __syncthreads();
      const int live = s_live;
      if (keep_list(live)) {
        keep_range(flags, 0, n, map_out, compose, c0_in, c0_out, c1_in, c1_out);
        if (t == 0) {
          *count_out = n;
          if (live_out) *live_out = live;
        }
        return;
      }
    }
    const int placed = compact_range<kCompactSub>(flags, 0, n, 0, map_out, compose, c0_in, c0_out,
                                                  c1_in, c1_out, s_cnt, s_off, &s_tot);
    if (t == 0) {
      *count_out = placed;
      if (live_out) *live_out = placed;
    }
    return;
  }
  const int64_t base = (int64_t)blockIdx.x * kCompactTile;
  if (base >= n) return;
  // output slots start after the earlier tiles' columns; the total decides whether the list
  // is kept
  const int ntiles = (int)((n + kCompactTile - 1) / kCompactTile);
  if (t == 0) {
    s_base = 0;
    s_live = 0;
  }
  __syncthreads();
  int part = 0, all = 0;
  for (int j = t; j < ntiles; j += 1024) {
    const int v = tile_cnt[j];
    all += v;
    if (j < (int)blockIdx.x) part += v;
  }
  for (int d = kWave / 2; d > 0; d >>= 1) {
    part += __shfl_xor(part, d, kWave);
    all += __shfl_xor(all, d, kWave);
  }
  if (lane == 0 && part) atomicAdd(&s_base, part);
  if (lane == 0 && all) atomicAdd(&s_live, all);
  __syncthreads();
  const int end = (int)min<int64_t>(base + kCompactTile, n);
  const int live = s_live;
  if (keep_list(live)) {
    keep_range(flags, (int)base, end, map_out, compose, c0_in, c0_out, c1_in, c1_out);
    if (end == n && t == 0) {
      *count_out = n;
      if (live_out) *live_out = live;
    }
    return;
  }
  const int placed = compact_range<kCompactTileSub>(flags, (int)base, end, s_base, map_out, compose,
                                                    c0_in, c0_out, c1_in, c1_out, s_cnt, s_off,
                                                    &s_tot);
  if (end == n && t == 0) {
    *count_out = placed;
    if (live_out) *live_out = placed;
  }
}

// Non-empty delta rows of the compact post layout: workgroup i counts the set flags of ring
// slot (j0 + i) % kPostRing over the list's *count_ptr positions into nerow[i] and clears
// them, except in the last slot of the range, which the compaction that follows consumes.
__global__ __launch_bounds__(1024) void ring_count_kernel(int32_t* __restrict__ ring,
                                                         int64_t ring_ld, int32_t j0,
                                                         int32_t nslots,
                                                         const int32_t* __restrict__ count_ptr,
                                                         int32_t* __restrict__ nerow) {
  __shared__ int32_t s_sum;
  const int i = blockIdx.x;
  int32_t* __restrict__ f = ring + (int64_t)((j0 + i) % kPostRing) * ring_ld;
  const bool clear = i + 1 < nslots;
  const int n = *count_ptr;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  int c = 0;
  for (int q = threadIdx.x; q < n; q += 1024) {
    const int32_t v = f[q];
    c += v != 0;
    if (clear && v) f[q] = 0;
  }
  for (int d = kWave / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, kWave);
  if (lane_id() == 0 && c) atomicAdd(&s_sum, c);
  __syncthreads();
  if (threadIdx.x == 0) nerow[i] = s_sum;
}

__global__ void fill_i32_kernel(int32_t* __restrict__ p, int32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void iota_i32_kernel(int32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}
__global__ void fill_f32_kernel(float* __restrict__ p, float v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Nonzero entries of the first B columns of rows x ld feature-major activations: one
// column range of 64 per lane group, rows strided over the grid; wavefront sums through
// cross-lane adds, one atomic per wavefront.
__global__ __launch_bounds__(256) void count_nonzero_kernel(const float* __restrict__ y,
                                                            int64_t rows, int64_t ld, int64_t B,
                                                            unsigned long long* __restrict__ out) {
  unsigned int c = 0;
  const int64_t tiles = (B + 63) / 64;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  for (int64_t i = w; i < rows * tiles; i += waves) {
    const int64_t n = i / tiles, pos = (i % tiles) * 64 + lane;
    if (pos < B && y[n * ld + pos] != 0.f) ++c;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if (lane == 0 && c) atomicAdd(out, (unsigned long long)c);
}

// 64 x 64 tile transpose through LDS (65-float rows avoid bank conflicts).
// in: [rows][cols] row-major, out[c * ldo + r].
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in,
                                                         int64_t rows, int64_t cols,
                                                         int64_t ldi, float* __restrict__ out,
                                                         int64_t ldo,
                                                         int32_t* __restrict__ nonfinite) {
  __shared__ float tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  bool bad = false;
  for (int k = ty; k < 64; k += 4) {
    const int64_t r = r0 + k, c = c0 + tx;
    const float v = (r < rows && c < cols) ? in[r * ldi + c] : 0.f;
    bad |= !__builtin_isfinite(v);
    tile[k][tx] = v;
  }
  if (nonfinite && bad) *nonfinite = 1;
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const int64_t c = c0 + k, r = r0 + tx;
    if (c < cols && r < rows) out[c * ldo + r] = tile[tx][k];
  }
}

__global__ __launch_bounds__(256) void scatter_cols_kernel(
    const float* __restrict__ src, int64_t ldsrc, const int32_t* __restrict__ map,
    const int32_t* __restrict__ dstcol, const int32_t* __restrict__ count_ptr, int32_t N,
    float* __restrict__ dst, int64_t lddst) {
  const int64_t count = *count_ptr;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int64_t s = map ? (int64_t)map[j] : j;
  const int64_t d = dstcol[j];
  const int n0 = blockIdx.y * 64, n1 = min(N, n0 + 64);
  for (int n = n0; n < n1; ++n) dst[(int64_t)n * lddst + d] = src[(int64_t)n * ldsrc + s];
}

// Binary blob images, one column per thread: pixel n of sample b is on iff it lies in a
// kept disc of the sample's prototype shifted by (dx, dy).
__global__ __launch_bounds__(256) void images_kernel(const int32_t* __restrict__ params,
                                                      const int32_t* __restrict__ blobs,
                                                      const int32_t* __restrict__ proto,
                                                      int32_t W, int32_t H, int64_t B,
                                                      int32_t N, float* __restrict__ Y,
                                                      int64_t ld) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int cls = params[b * 4 + 0], dx = params[b * 4 + 1], dy = params[b * 4 + 2],
            keep = params[b * 4 + 3];
  const int first = proto[cls], nb = proto[10 + cls];
  const int n0 = blockIdx.y * 256, n1 = min(N, n0 + 256);
  for (int n = n0; n < n1; ++n) {
    const int x = n % W, y = n / W;
    float v = 0.f;
    if (y < H) {
      for (int q = 0; q < nb; ++q) {
        if (!((keep >> q) & 1)) continue;
        const int32_t* bl = blobs + (first + q) * 4;
        const int ddx = x - bl[0] - dx, ddy = y - bl[1] - dy;
        if (ddx * ddx + ddy * ddy <= bl[2] * bl[2]) {
          v = 1.f;
          break;
        }
      }
    }
    Y[(int64_t)n * ld + b] = v;
  }
}

__global__ __launch_bounds__(256) void row_any_kernel(const float* __restrict__ Y, int64_t ld,
                                                       const int32_t* __restrict__ count_ptr,
                                                       int32_t N, int32_t* __restrict__ flags) {
  const int64_t count = *count_ptr;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count) return;
  const int n0 = blockIdx.y * 256, n1 = min(N, n0 + 256);
  bool any = false;
  for (int n = n0; n < n1; ++n) any |= Y[(int64_t)n * ld + b] != 0.f;
  if (any) flags[b] = 1;
}

__global__ void cat_from_list_kernel(const int32_t* __restrict__ orig,
                                     const int32_t* __restrict__ count_ptr,
                                     int32_t* __restrict__ cat) {
  const int64_t count = *count_ptr;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count;
       j += (int64_t)gridDim.x * blockDim.x)
    cat[orig[j]] = 1;
}
__global__ void cat_from_flags_kernel(int32_t* __restrict__ flags,
                                      const int32_t* __restrict__ count_ptr,
                                      int32_t* __restrict__ cat) {
  const int64_t count = *count_ptr;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count;
       j += (int64_t)gridDim.x * blockDim.x) {
    cat[j] = flags[j] != 0;
    flags[j] = 0;
  }
}

// ---- exact clustering ----------------------------------------------------------------
// Pool row r of segment s (its r-th live row, r < min(pool, live in s)) gets `segments`
// sums over consecutive blocks of N/segments neurons (the intended segments of
// src/SDGC/SNICIT/kernel.hpp:138-153, which accumulates them with shared-memory atomics in no
// fixed order).  Fixed order here: each segment splits into S2 = min(32, seglen) contiguous
// sub-chunks [c*seglen/S2, (c+1)*seglen/S2), each summed in ascending neuron order, then the
// S2 partial sums are added in ascending c.  Block (pool row, batch segment), thread (g, c).
__global__ __launch_bounds__(1024) void seg_sums_kernel(const float* __restrict__ Y, int64_t ld,
                                                         const int32_t* __restrict__ map,
                                                         const int32_t* __restrict__ seglo,
                                                         int32_t pool, int32_t segments,
                                                         int32_t N, float* __restrict__ F) {
  __shared__ float part[32 * 32];
  const int r = blockIdx.x, sg = blockIdx.y;
  const int lo = seglo[sg];
  const int P = min(pool, seglo[sg + 1] - lo);
  const int seglen = N / segments;
  const int S2 = min(32, seglen);
  const int g = threadIdx.x / 32, c = threadIdx.x % 32;
  if (r < P && g < segments && c < S2) {
    const int64_t pos = lo + r;
    const int64_t col = map ? (int64_t)map[pos] : pos;
    const int n0 = g * seglen + c * seglen / S2, n1 = g * seglen + (c + 1) * seglen / S2;
    const float* __restrict__ y = Y + col;
    float sum = 0.f;
    int n = n0;
    for (; n + 16 <= n1; n += 16) {
      float v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = y[(int64_t)(n + q) * ld];
#pragma unroll
      for (int q = 0; q < 16; ++q) sum += v[q];
    }
    for (; n < n1; ++n) sum += y[(int64_t)n * ld];
    part[g * 32 + c] = sum;
  }
  __syncthreads();
  if (threadIdx.x < segments) {
    float sum = 0.f;
    if (r < P)
      for (int q = 0; q < S2; ++q) sum += part[threadIdx.x * 32 + q];
    F[((int64_t)sg * pool + r) * segments + threadIdx.x] = sum;
  }
}

// Alg. 1 sample pruning (src/SDGC/SNICIT/kernel.hpp:156-194): in pool order, every still
// present base c removes each other present row y whose segment sums all lie within eps
// of c's (count of |F_y - F_c| > eps below segments*eps).  Survivors are the centroids,
// listed in pool order as list positions in cent_slot[s*stride ...].  One wavefront per
// batch segment, lane = pool row.  kcols (zero on entry) becomes the number of centroid
// columns in use, s*stride + k_s of the last segment with centroids.
__global__ __launch_bounds__(64) void select_exact_kernel(const float* __restrict__ F,
                                                           const int32_t* __restrict__ seglo,
                                                           int32_t nseg, int32_t pool,
                                                           int32_t segments, float cut,
                                                           double limit, int32_t stride,
                                                           int32_t* __restrict__ cent_slot,
                                                           int32_t* __restrict__ kseg,
                                                           int32_t* __restrict__ kcols) {
  const int lane = threadIdx.x, sg = blockIdx.x;
  const int lo = seglo[sg];
  const int P = min(pool, seglo[sg + 1] - lo);
  const float* __restrict__ Fs = F + (int64_t)sg * pool * segments;
  bool alive = lane < P;
  for (int c = 0; c < P; ++c) {
    const uint64_t am = __ballot(alive);
    if (!((am >> c) & 1ull)) continue;
    if (alive && lane != c) {
      int cnt = 0;
      for (int g = 0; g < segments; ++g)
        cnt += fabsf(Fs[lane * segments + g] - Fs[c * segments + g]) > cut;
      if ((double)cnt < limit) alive = false;
    }
  }
  const uint64_t am = __ballot(alive);
  if (alive) cent_slot[sg * stride + lane_prefix(am)] = lo + lane;
  if (lane == 0) {
    const int k = __popcll(am);
    kseg[sg] = k;
    if (k > 0) atomicMax(kcols, sg * stride + k);
  }
}

// Centroid column c belongs to segment c / stride and is in use when c % stride < kseg.
__global__ __launch_bounds__(256) void gather_centroids_kernel(
    const float* __restrict__ Y, int64_t ld, const int32_t* __restrict__ map,
    const int32_t* __restrict__ cent_slot, const int32_t* __restrict__ kseg, int32_t stride,
    int32_t N, float* __restrict__ C, int32_t kcap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N * kcap) return;
  const int64_t n = i / kcap;
  const int c = (int)(i % kcap);
  float v = 0.f;
  if (c % stride < kseg[c / stride]) {
    const int64_t slot = cent_slot[c];
    const int64_t col = map ? (int64_t)map[slot] : slot;
    v = Y[n * ld + col];
  }
  C[i] = v;
}

// Nearest centroid (src/SDGC/SNICIT/kernel.hpp:8-69): m(i) = argmin_c #{n : Y_i[n] != Y_c[n]}
// over the centroids of i's batch segment, first minimum in centroid order; centroids map to
// themselves.  flag = non-centroid with a nonzero distance (its delta row is non-empty).
// The distances are split over blocks of neurons (grid y) and summed with integer atomics
// into dist[i][local centroid].  A wavefront walks the segments its lanes cover one at a time
// (almost always one), so the centroid values stay uniform (scalar loads); eight running
// counts per pass.
__global__ __launch_bounds__(256) void assign_dist_kernel(
    const float* __restrict__ Y, int64_t ld, const int32_t* __restrict__ map,
    const int32_t* __restrict__ seglo, int32_t nseg, const float* __restrict__ C, int32_t kcap,
    int32_t stride, const int32_t* __restrict__ kseg, int32_t N, int32_t nchunk,
    int32_t* __restrict__ dist) {
  const int64_t count = seglo[nseg];
  if ((int64_t)blockIdx.x * blockDim.x >= count) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t col = map ? (int64_t)map[i] : i;
  const int sg = nseg == 1 ? 0 : segment_of(seglo, nseg, i);
  const int n0 = (int)((int64_t)N * blockIdx.y / nchunk);
  const int n1 = (int)((int64_t)N * (blockIdx.y + 1) / nchunk);
  uint64_t todo = __ballot(true);
  while (todo) {
    const int lead = __ffsll((unsigned long long)todo) - 1;
    const int s0 = __builtin_amdgcn_readfirstlane(__shfl(sg, lead));
    const bool mine = sg == s0;
    todo &= ~__ballot(mine);
    const int k = kseg[s0];
    const float* __restrict__ Cs = C + (int64_t)s0 * stride;
    // centroids cb..cb+7 of the segment: one block of eight per grid z index
    for (int cb = 8 * (int)blockIdx.z; cb < k; cb += 8 * (int)gridDim.z) {
      int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int n = n0; n < n1; ++n) {
        const float y = Y[(int64_t)n * ld + col];
        const float* __restrict__ cr = Cs + (int64_t)n * kcap + cb;
#pragma unroll
        for (int q = 0; q < 8; ++q) cnt[q] += (y != cr[q]);
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (mine && cb + q < k && cnt[q]) atomicAdd(&dist[i * kMaxCentroidsExact + cb + q], cnt[q]);
    }
  }
}

__global__ __launch_bounds__(256) void assign_final_kernel(
    const int32_t* __restrict__ seglo, int32_t nseg, const int32_t* __restrict__ kseg,
    int32_t stride, const int32_t* __restrict__ cent_slot, int32_t* __restrict__ dist,
    int32_t* __restrict__ flags, int32_t* __restrict__ mcar, const int32_t* __restrict__ orig,
    int32_t* __restrict__ rec) {
  const int64_t count = seglo[nseg];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int sg = segment_of(seglo, nseg, i);
  const int k = kseg[sg];
  const int32_t* __restrict__ cs = cent_slot + sg * stride;
  int best = 0x7fffffff, bestc = 0;
  for (int c = 0; c < k; ++c) {
    const int d = dist[i * kMaxCentroidsExact + c];
    dist[i * kMaxCentroidsExact + c] = 0;  // ready for the next run
    if (d < best) {
      best = d;
      bestc = c;
    }
  }
  bool is_cent = false;
  for (int c = 0; c < k; ++c)
    if (cs[c] == i) {
      is_cent = true;
      bestc = c;
    }
  const int32_t m = sg * stride + bestc;
  mcar[i] = m;
  rec[orig[i]] = m;
  if (!is_cent && best > 0) flags[i] = 1;
}

__global__ __launch_bounds__(256) void make_delta_kernel(
    const float* __restrict__ Y, int64_t ld, const int32_t* __restrict__ map,
    const float* __restrict__ C, int32_t kcap, const int32_t* __restrict__ m,
    const int32_t* __restrict__ count_ptr, int32_t N, float* __restrict__ D, int64_t ldD_arg,
    int32_t compact) {
  const int64_t count = *count_ptr;
  const int64_t ldD = compact ? compact_ld(count) : ldD_arg;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int64_t col = map[j];
  const int c = m[j];
  const int n0 = blockIdx.y * 64, n1 = min(N, n0 + 64);
#pragma unroll 16
  for (int n = n0; n < n1; ++n)
    D[(int64_t)n * ldD + j] = Y[(int64_t)n * ld + col] - C[(int64_t)n * kcap + c];
}

__global__ __launch_bounds__(256) void centroid_cats_kernel(const float* __restrict__ C,
                                                             int32_t kcap,
                                                             const int32_t* __restrict__ k_ptr,
                                                             int32_t N,
                                                             int32_t* __restrict__ catc) {
  const int c = blockIdx.x;
  if (c >= *k_ptr) return;
  int any = 0;
  for (int n = threadIdx.x; n < N; n += blockDim.x) any |= C[(int64_t)n * kcap + c] > 0.f;
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) catc[c] = any ? 1 : 0;
}

__global__ void recover_base_kernel(const int32_t* __restrict__ rec,
                                    const int32_t* __restrict__ B_ptr,
                                    const int32_t* __restrict__ catc, int32_t* __restrict__ cat) {
  const int64_t B = *B_ptr;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < B;
       b += (int64_t)gridDim.x * blockDim.x) {
    const int r = rec[b];
    cat[b] = r < 0 ? 0 : catc[r];
  }
}

// active(i) = any(Yhat_i + Yhat_m(i) > 0) over a 256-neuron slice per block row.
__global__ __launch_bounds__(256) void recover_rows_kernel(
    const float* __restrict__ D, int64_t ldD_arg, const int32_t* __restrict__ ldD_cnt,
    const int32_t* __restrict__ map, const int32_t* __restrict__ m,
    const int32_t* __restrict__ count_ptr, const float* __restrict__ C, int32_t kcap, int32_t N,
    int32_t* __restrict__ act) {
  const int64_t count = *count_ptr;
  const int64_t ldD = ldD_cnt ? compact_ld(*ldD_cnt) : ldD_arg;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int64_t col = map[j];
  const int c = m[j];
  const int n0 = blockIdx.y * 64, n1 = min(N, n0 + 64);
  bool any = false;
#pragma unroll 16
  for (int n = n0; n < n1; ++n)
    any |= (D[(int64_t)n * ldD + col] + C[(int64_t)n * kcap + c]) > 0.f;
  if (any) act[j] = 1;
}

__global__ void recover_write_kernel(const int32_t* __restrict__ orig,
                                     const int32_t* __restrict__ count_ptr,
                                     int32_t* __restrict__ act, int32_t* __restrict__ cat) {
  const int64_t count = *count_ptr;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count;
       j += (int64_t)gridDim.x * blockDim.x) {
    cat[orig[j]] = act[j];
    act[j] = 0;
  }
}

__global__ __launch_bounds__(256) void expand_centroids_kernel(
    const int32_t* __restrict__ rec, const int32_t* __restrict__ B_ptr,
    const float* __restrict__ C, int32_t kcap, int32_t N, float* __restrict__ Y, int64_t ld) {
  const int64_t B = *B_ptr;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int r = rec[b];
  const int n0 = blockIdx.y * 64, n1 = min(N, n0 + 64);
  for (int n = n0; n < n1; ++n) Y[(int64_t)n * ld + b] = r < 0 ? 0.f : C[(int64_t)n * kcap + r];
}

__global__ __launch_bounds__(256) void scatter_delta_kernel(
    const float* __restrict__ D, int64_t ldD_arg, const int32_t* __restrict__ ldD_cnt,
    const int32_t* __restrict__ map, const int32_t* __restrict__ m,
    const int32_t* __restrict__ orig, const int32_t* __restrict__ count_ptr,
    const float* __restrict__ C, int32_t kcap, int32_t N, float* __restrict__ Y, int64_t ld) {
  const int64_t count = *count_ptr;
  const int64_t ldD = ldD_cnt ? compact_ld(*ldD_cnt) : ldD_arg;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int64_t col = map[j], dst = orig[j];
  const int c = m[j];
  const int n0 = blockIdx.y * 64, n1 = min(N, n0 + 64);
  for (int n = n0; n < n1; ++n)
    Y[(int64_t)n * ld + dst] = D[(int64_t)n * ldD + col] + C[(int64_t)n * kcap + c];
}

// ---- dense layers on the matrix cores -----------------------------------------------
// Dense input layer (src/beyond/general.hpp:5-33) as a GEMM on the matrix cores.  A
// wavefront owns RB 32-row blocks x CB 64-column blocks; step k multiplies the weights of
// input k for its rows (A: lane l holds row l mod 32 of a block, both halves) with input k
// of its columns (B: lane l holds column l of a block) into RB*CB accumulators, k
// ascending.  With K = 1 per step every output element takes one fused multiply-add per
// input, so the result is exactly the chain acc = fma(w[K-1], x[K-1], ... fma(w[0], x[0],
// 0)) of the vector kernels and the oracle (D layout as in panel_mfma_kernel).  Operands
// come from the transposed weights wt[k][row] and the input rows, U steps ahead in a second
// register set.  The four wavefronts of a workgroup take four row tiles of one column
// tile, so the tile's inputs come from L1 after the first.  Columns past the live count read
// the last live column; their outputs are dropped.  With at most `narrow` live columns every
// thread of the grid takes (row, column) chains over the row-major weights instead.  A
// launch handles live counts in [lo, hi) and returns at once otherwise, so two tile forms
// can be launched back to back and the live count picks one on the device.
// Input steps per register set: 16 for the one-accumulator form (its chain is bound by
// the latency of its loads), 8 for the four-accumulator form (bound by the matrix cores).
template <int RB, int CB>
struct DenseSteps {
  static constexpr int U = RB * CB == 1 ? 16 : 8;
};

template <int EPI, int RB, int CB>
__global__ __launch_bounds__(256) void dense_mfma_kernel(
    const float* __restrict__ wt, int32_t ldw, const float* __restrict__ vals, int32_t K,
    int32_t rows, const float* __restrict__ bias_vec, float bias_s, float clip,
    const float* __restrict__ yin, int64_t ldin, float* __restrict__ yout, int64_t ldout,
    const int32_t* __restrict__ count_ptr,
    int32_t narrow, int64_t lo, int64_t hi) {
  constexpr int U = DenseSteps<RB, CB>::U;
  const int64_t count = *count_ptr;
  if (count < lo || count >= hi) return;
  if (count <= narrow) {
    ell_pair_chains<EPI>(nullptr, vals, K, rows, bias_vec, bias_s, clip, yin, ldin, yout, ldout,
                         count, nullptr, nullptr, nullptr, 0, nullptr, 0.f,
                         (int64_t)blockIdx.x * 256 + threadIdx.x, (int64_t)gridDim.x * 256);
    return;
  }
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int half = lane >> 5, j = lane & 31;
  const int row_tiles = ldw / (32 * RB);  // ldw: rows padded to whole tiles (zero weights)
  const int64_t items = (int64_t)row_tiles * ((count + 64 * CB - 1) / (64 * CB));
  for (int64_t it = (int64_t)blockIdx.x * 4 + wv; it < items; it += (int64_t)gridDim.x * 4) {
    const int r0 = (int)(it % row_tiles) * 32 * RB;
    const int64_t c0 = it / row_tiles * 64 * CB;
    const float* __restrict__ ap = wt + r0 + j;
    const float* __restrict__ bp[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) bp[cb] = yin + min(c0 + 64 * cb + lane, count - 1);
    float a0[U][RB], b0[U][CB], a1[U][RB], b1[U][CB];
    floatx32 acc[RB][CB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int q = 0; q < 32; ++q) acc[rb][cb][q] = 0.f;
    auto load = [&](float (&a)[U][RB], float (&b)[U][CB], int k0) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) a[u][rb] = ap[(int64_t)(k0 + u) * ldw + 32 * rb];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) b[u][cb] = bp[cb][(int64_t)(k0 + u) * ldin];
      }
    };
    auto mma = [&](const float (&a)[U][RB], const float (&b)[U][CB]) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
          for (int cb = 0; cb < CB; ++cb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x1f32(a[u][rb], b[u][cb], acc[rb][cb],
                                                               0, 0, 0);
    };
    // whole pairs of register sets, then the last K mod 2U inputs one at a time (k ascending)
    const int Kmain = K - K % (2 * U);
    if (Kmain > 0) load(a0, b0, 0);
    for (int k0 = 0; k0 < Kmain; k0 += 2 * U) {
      load(a1, b1, k0 + U);
      mma(a0, b0);
      if (k0 + 2 * U < Kmain) load(a0, b0, k0 + 2 * U);
      mma(a1, b1);
    }
    // (K is a multiple of 16: with 2U = 16 there is no tail)
    for (int k = Kmain; 2 * U > 16 && k < K; ++k) {
      float a[RB], b[CB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) a[rb] = ap[(int64_t)k * ldw + 32 * rb];
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) b[cb] = bp[cb][(int64_t)k * ldin];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x1f32(a[rb], b[cb], acc[rb][cb], 0, 0, 0);
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      float bias[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int o = r0 + 32 * rb + 8 * (q >> 2) + 4 * half + (q & 3);
        bias[q] = bias_vec && o < rows ? bias_vec[o] : bias_s;
      }
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const int64_t pos = c0 + 64 * cb + 32 * b + j;
          if (pos >= count) continue;
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int o = r0 + 32 * rb + 8 * (q >> 2) + 4 * half + (q & 3);
            bool f = false;
            const float y = epilogue<EPI>(acc[rb][cb][16 * b + q], bias[q], clip, true, o, pos,
                                          0, nullptr, 0, 0.f, f);
            if (o < rows) yout[(int64_t)o * ldout + pos] = y;
          }
        }
    }
  }
}

// Workgroups along the column-tile axis of a (row group x column tile) grid: grid_target /
// groups, at most one per tile, and at most one per kMinTilesPerBlock tiles of the plan's
// capacity.  Launch sizes follow the capacity, not the live count, so a plan whose
// capacity is only a few tiles deep (a serving slot) would otherwise launch mostly
// workgroups that find no columns; each workgroup loops over its tiles.
constexpr int64_t kMinTilesPerBlock = 4;
inline int64_t grid_rows(int64_t grid_target, int64_t groups, int64_t tiles) {
  int64_t gy = std::max<int64_t>(1, grid_target / std::max<int64_t>(groups, 1));
  gy = std::min<int64_t>(gy, std::max<int64_t>(1, (tiles + kMinTilesPerBlock - 1) / kMinTilesPerBlock));
  return std::min<int64_t>(gy, 65535);
}

inline unsigned grid1(int64_t n, int block, int64_t cap = 65535) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

inline void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw Error(SNICIT_E_HIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
}

}  // namespace

double tol_decimal(float e) {
  char buf[32];
  for (int p = 1; p <= 9; ++p) {
    snprintf(buf, sizeof buf, "%.*g", p, (double)e);
    if (strtof(buf, nullptr) == e) return strtod(buf, nullptr);
  }
  return (double)e;
}

float tol_cut(float e) {
  const double d = tol_decimal(e);
  float c = (float)d;
  if ((double)c > d) c = nextafterf(c, -INFINITY);
  return c;
}

void launch_panel_spmm(const SpmmLaunch& a, hipStream_t s) {
  const Net& net = *a.net;
  const int l = a.layer;
  const int64_t P = net.panels(l);
  const int64_t tiles = (a.max_cols + kBlock - 1) / kBlock;
  // Enough workgroups to fill 256 CUs several times over; each loops over column tiles.
  const int64_t gy = grid_rows(a.shape.grid_target, P, tiles);
  SNICIT_REQUIRE(P <= 2147483647LL, "too many panels");
  dim3 grid((unsigned)P, (unsigned)gy);
  const int32_t* pcols = a.la.pcols;
  const int32_t* pbeg = a.la.pbeg;
  const int32_t* outs = a.la.outs;
  const float* w = a.la.w;
  const float* bv = net.bias_vec(l);
  const float bs = net.bias_scalar(l);
  const bool mfma = net.mfma[l] != 0;
  // panels whose rows all repeat one weight row need one chain per column: the vector
  // kernel runs it; otherwise plain layers take the 32 chains on the matrix cores
  if (!mfma && a.epilogue == EPI_PLAIN && net.plain_mfma && !net.dup_rows[l]) {
    hipLaunchKernelGGL(panel_plain_mfma_kernel, grid, dim3(kBlock), 0, s, pcols, pbeg, outs, w,
                       bv, bs, net.clip, a.yin, a.ldin, a.yout, a.ldout, a.count, a.map, a.flags,
                       a.shape.narrow);
    check_launch();
    return;
  }
#define SNICIT_SPMM(E)                                                                         \
  do {                                                                                         \
    if (mfma)                                                                                  \
      hipLaunchKernelGGL(panel_mfma_kernel<E>, grid, dim3(kBlock), 0, s, pcols, pbeg, outs, w, \
                         bv, bs, net.clip, a.yin, a.ldin, a.yout, a.ldout, a.count, a.map,     \
                         a.flags, a.zc, a.kcap, a.mcent, a.eps_post);                          \
    else                                                                                       \
      hipLaunchKernelGGL(panel_spmm_kernel<E>, grid, dim3(kBlock), 0, s, pcols, pbeg, outs, w, \
                         bv, bs, net.clip, a.yin, a.ldin, a.yout, a.ldout, a.count, a.map,     \
                         a.flags, a.zc, a.kcap, a.mcent, a.eps_post, a.shape.narrow,           \
                         a.la.psame);                                                          \
  } while (0)
  if (a.epilogue == EPI_PLAIN && !mfma && (a.shape.plain_tiles == 2 || a.shape.plain_tiles == 4)) {
    if (a.shape.plain_tiles == 2)
      hipLaunchKernelGGL(panel_spmm_tiles_kernel<2>, grid, dim3(kBlock), 0, s, pcols, pbeg, outs,
                         w, bv, bs, net.clip, a.yin, a.ldin, a.yout, a.ldout, a.count, a.map,
                         a.flags, a.shape.narrow, a.la.psame);
    else
      hipLaunchKernelGGL(panel_spmm_tiles_kernel<4>, grid, dim3(kBlock), 0, s, pcols, pbeg, outs,
                         w, bv, bs, net.clip, a.yin, a.ldin, a.yout, a.ldout, a.count, a.map,
                         a.flags, a.shape.narrow, a.la.psame);
    check_launch();
    return;
  }
  switch (a.epilogue) {
    case EPI_PLAIN: SNICIT_SPMM(EPI_PLAIN); break;
    case EPI_CENTROID: SNICIT_SPMM(EPI_CENTROID); break;
    case EPI_DELTA: SNICIT_SPMM(EPI_DELTA); break;
    case EPI_DELTA_LOSSY: SNICIT_SPMM(EPI_DELTA_LOSSY); break;
    default: throw Error(SNICIT_E_INVALID, "bad epilogue");
  }
#undef SNICIT_SPMM
  check_launch();
}

bool post_layer_fused(const Net& net, int l) { return net.mfma[l] == 0; }

void launch_panel_centroid(const Net& net, int l, const LayerArrays& la, const float* cin,
                           float* cout, float* zc, int32_t kcap, const int32_t* k,
                           hipStream_t s) {
  SNICIT_REQUIRE(net.mfma[l] == 0, "centroid kernel needs row-major weights");
  SNICIT_REQUIRE(kcap == kMaxCentroidsExact, "centroid kernel expects 32 centroid slots");
  const int32_t* pcols = la.pcols;
  const int32_t* pbeg = la.pbeg;
  const int32_t* outs = la.outs;
  const float* w = la.w;
  hipLaunchKernelGGL(panel_centroid_kernel, dim3((unsigned)net.panels(l)), dim3(kBlock), 0, s,
                     pcols, pbeg, outs, w, net.bias_vec(l), net.bias_scalar(l), net.clip, cin,
                     cout, zc, kcap, k);
  check_launch();
}

void launch_panel_post(const Net& net, int l, const LayerArrays& la, bool lossy,
                       const float* cin, float* cout, int32_t kcap, const int32_t* k,
                       const float* din, int64_t ldin, const int32_t* ldin_cnt, float* dout,
                       int64_t ldout, bool compact_out, const int32_t* count, int32_t* nlist_next,
                       const int32_t* map, const int32_t* mcent, int32_t* flags, float eps_post,
                       int64_t max_cols, const LaunchShape& shape, hipStream_t s,
                       bool ccompact, const float* zchain, const NextChain& nx) {
  SNICIT_REQUIRE(net.mfma[l] == 0, "post kernel needs row-major weights");
  SNICIT_REQUIRE(!zchain || (net.few_rows[l] && shape.post_form >= 1),
                 "chain records need the wavefront-local post kernel");
  SNICIT_REQUIRE(kcap == kMaxCentroidsExact, "post kernel expects 32 centroid slots");
  const int64_t P = net.panels(l);
  const int64_t tiles = (max_cols + kBlock - 1) / kBlock;  // 256-column tiles per panel
  const int64_t gy = grid_rows(shape.post_grid_target, P, tiles);
  dim3 grid((unsigned)P, (unsigned)gy);
  const int32_t* pcols = la.pcols;
  const int32_t* pbeg = la.pbeg;
  const int32_t* outs = la.outs;
  const float* w = la.w;
  // panels of at most two distinct weight rows: the wavefront-local form
  SNICIT_REQUIRE(!nx.pbeg || (zchain && net.few_rows[l] && shape.post_form >= 1),
                 "the next layer's chains need this layer's chain records");
  if (net.few_rows[l] && shape.post_form >= 1) {
    if (nx.pbeg) grid.x = (unsigned)std::max<int64_t>(P, nx.P);
#define SNICIT_POST_WAVE(LO, TWO)                                                               \
  hipLaunchKernelGGL((panel_post_wave_kernel<LO, TWO>), grid, dim3(nx.pbeg ? kBlock + kWave : kBlock), \
                     0, s, pcols, pbeg,                                                           \
                     outs, w, net.bias_vec(l), net.bias_scalar(l), net.clip, zchain, cin, cout,  \
                     kcap, k, din, ldin, ldin_cnt, dout, ldout, compact_out ? 1 : 0, count,      \
                     nlist_next, map, mcent, flags, eps_post, la.psame, ccompact ? 1 : 0, P,  \
                     nx)
    const bool two = !net.dup_rows[l];
    if (lossy) {
      if (two) SNICIT_POST_WAVE(true, true); else SNICIT_POST_WAVE(true, false);
    } else {
      if (two) SNICIT_POST_WAVE(false, true); else SNICIT_POST_WAVE(false, false);
    }
#undef SNICIT_POST_WAVE
    check_launch();
    return;
  }
  if (lossy)
    hipLaunchKernelGGL(panel_post_kernel<true>, grid, dim3(kBlock), 0, s, pcols, pbeg, outs, w,
                       net.bias_vec(l), net.bias_scalar(l), net.clip, cin, cout, kcap, k, din,
                       ldin, ldin_cnt, dout, ldout, compact_out ? 1 : 0, count, nlist_next, map,
                       mcent, flags, eps_post, shape.post_narrow, la.psame, ccompact ? 1 : 0);
  else
    hipLaunchKernelGGL(panel_post_kernel<false>, grid, dim3(kBlock), 0, s, pcols, pbeg, outs, w,
                       net.bias_vec(l), net.bias_scalar(l), net.clip, cin, cout, kcap, k, din,
                       ldin, ldin_cnt, dout, ldout, compact_out ? 1 : 0, count, nlist_next, map,
                       mcent, flags, eps_post, shape.post_narrow, la.psame, ccompact ? 1 : 0);
  check_launch();
}

void launch_post_chain(const Net& net, int l, const LayerArrays& la, const float* cin,
                       float* cout, int32_t kcap, const int32_t* k, bool ccompact, float* zchain,
                       hipStream_t s) {
  SNICIT_REQUIRE(net.mfma[l] == 0 && net.few_rows[l], "chain records need panels of at most "
                 "two distinct row-major weight rows");
  SNICIT_REQUIRE(kcap == kMaxCentroidsExact, "chain records expect 32 centroid slots");
  const int64_t P = net.panels(l);
  const int64_t per = kBlock / kWave;
  hipLaunchKernelGGL(post_chain_kernel, dim3((unsigned)((P + per - 1) / per)), dim3(kBlock), 0, s,
                     la.pcols, la.pbeg, la.outs, la.w, la.psame, net.bias_vec(l),
                     net.bias_scalar(l), net.clip, cin, cout, kcap, k, ccompact ? 1 : 0, zchain,
                     P);
  check_launch();
}

void launch_centroid_restride(const float* src, bool src_compact, float* dst, bool dst_compact,
                              int32_t kcap, const int32_t* k, int64_t N, hipStream_t s) {
  hipLaunchKernelGGL(centroid_restride_kernel, dim3(grid1(N * kcap, 256, 4096)), dim3(256), 0, s,
                     src, src_compact ? 1 : 0, dst, dst_compact ? 1 : 0, kcap, k, N);
  check_launch();
}

// Columns per LDS tile of the staged kernel: the largest of 32/16/8/4 whose tile (row stride
// TC+4 floats) fits the LDS budget, 0 when none does (or the budget is 0: direct gathers).
int ell_tile_cols(int64_t in_rows, int64_t lds_bytes) {
  for (int tc : {32, 16, 8, 4})
    if (in_rows * (tc + 4) * 4 <= lds_bytes) return tc;
  return 0;
}

template <int TC, bool RPT>
void launch_ell_lds(const EllLaunch& a, hipStream_t s) {
  const int64_t tiles = (a.max_cols + TC - 1) / TC;
  const int64_t groups = RPT ? (a.rows + 255) / 256 : 1;
  int64_t gx = std::max<int64_t>(1, a.shape.grid_target / std::max<int64_t>(groups, 1));
  gx = std::min<int64_t>(gx, std::max<int64_t>(tiles, 1));
  gx = std::min<int64_t>(gx, 1 << 30);
  dim3 grid((unsigned)gx, (unsigned)groups);
  const size_t lds = (size_t)a.in_rows * (TC + 4) * 4;
#define SNICIT_ELL_LDS(E)                                                                      \
  hipLaunchKernelGGL((ell_lds_kernel<E, TC, RPT>), grid, dim3(256), lds, s, a.cols, a.vals,    \
                     a.K, a.rows, a.in_rows, a.bias_vec, a.bias_s, a.clip, a.yin, a.ldin,       \
                     a.yout, a.ldout, a.count, a.map, a.flags, a.zc, a.kcap, a.mcent,           \
                     a.eps_post)
  switch (a.epilogue) {
    case EPI_PLAIN: SNICIT_ELL_LDS(EPI_PLAIN); break;
    case EPI_CENTROID: SNICIT_ELL_LDS(EPI_CENTROID); break;
    case EPI_DELTA: SNICIT_ELL_LDS(EPI_DELTA); break;
    case EPI_DELTA_LOSSY: SNICIT_ELL_LDS(EPI_DELTA_LOSSY); break;
    case EPI_LINEAR: SNICIT_ELL_LDS(EPI_LINEAR); break;
    default: throw Error(SNICIT_E_INVALID, "bad epilogue");
  }
#undef SNICIT_ELL_LDS
  check_launch();
}

template <int TC>
void launch_ell_lds_form(const EllLaunch& a, hipStream_t s) {
  // few rows: thread per (column, row slot); otherwise thread per row
  if (a.rows < 128) launch_ell_lds<TC, false>(a, s);
  else launch_ell_lds<TC, true>(a, s);
}

// CU count of the current device (cached per device)
int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  SNICIT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) dev = 0;
  int cus = cache[dev].load(std::memory_order_relaxed);
  if (cus <= 0) {
    SNICIT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cache[dev].store(cus, std::memory_order_relaxed);
  }
  return cus;
}

void launch_ell_chunk(const EllLaunch& a, int32_t rc, hipStream_t s) {
  constexpr int TC = 64;
  const int64_t tiles = (a.max_cols + TC - 1) / TC;
  const int64_t groups = (a.rows + 255) / 256;
  // enough workgroups for the (row, column) chains when the count may be narrow
  const int64_t pairs = a.rows * std::min<int64_t>(a.max_cols, a.shape.narrow);
  const int64_t gy_pairs = std::min<int64_t>((pairs + 255) / 256, (int64_t)8 * device_cus());
  const int64_t gy_tiles = grid_rows(a.shape.grid_target, groups, tiles);
  const int64_t gy = std::min<int64_t>(65535, std::max(gy_tiles, (gy_pairs + groups - 1) / groups));
  dim3 grid((unsigned)groups, (unsigned)gy);
  const size_t lds = (size_t)rc * (TC + 4) * 4;
#define SNICIT_ELL_CHUNK(E)                                                                    \
  hipLaunchKernelGGL(ell_chunk_kernel<E>, grid, dim3(256), lds, s, a.cols, a.vals, a.K, a.rows, \
                     a.in_rows, rc, a.bias_vec, a.bias_s, a.clip, a.yin, a.ldin, a.yout,        \
                     a.ldout, a.count, a.map, a.flags, a.zc, a.kcap, a.mcent, a.eps_post,       \
                     a.shape.narrow)
  switch (a.epilogue) {
    case EPI_PLAIN: SNICIT_ELL_CHUNK(EPI_PLAIN); break;
    case EPI_CENTROID: SNICIT_ELL_CHUNK(EPI_CENTROID); break;
    case EPI_DELTA: SNICIT_ELL_CHUNK(EPI_DELTA); break;
    case EPI_DELTA_LOSSY: SNICIT_ELL_CHUNK(EPI_DELTA_LOSSY); break;
    case EPI_LINEAR: SNICIT_ELL_CHUNK(EPI_LINEAR); break;
    default: throw Error(SNICIT_E_INVALID, "bad epilogue");
  }
#undef SNICIT_ELL_CHUNK
  check_launch();
}


template <int W>
void launch_ell_wave_w(const EllLaunch& a, hipStream_t s) {
  const int cus = device_cus();
  // LDS per workgroup: two chunks of a 64*W column tile (32 KB for W = 1, 128 KB for
  // W = 4): two workgroups per CU, or one
  const int per_cu = W == 1 ? 2 : 1;
  const int64_t tiles = (a.max_cols + 64 * W - 1) / (64 * W);
  const int64_t most = std::max<int64_t>(1, tiles * ((a.rows + kEllWaves - 1) / kEllWaves));
  const int64_t resident = (int64_t)per_cu * std::max(cus, 1);
  const unsigned grid = (unsigned)std::min<int64_t>(resident, most);
  const size_t lds = (size_t)2 * kEllChunk * 64 * W * sizeof(float);
#define SNICIT_ELL_WAVE(E)                                                                     \
  hipLaunchKernelGGL((ell_wave_kernel<E, W>), dim3(grid), dim3(kEllWaves * 64), lds, s, a.ent, \
                     a.rows, a.in_rows, a.rng, a.bias_vec, a.bias_s, a.clip, a.yin, a.ldin,     \
                     a.yout, a.ldout, a.count, a.map, a.flags, a.zc, a.kcap, a.mcent,           \
                     a.eps_post, (int32_t)resident)
  switch (a.epilogue) {
    case EPI_PLAIN: SNICIT_ELL_WAVE(EPI_PLAIN); break;
    case EPI_CENTROID: SNICIT_ELL_WAVE(EPI_CENTROID); break;
    case EPI_DELTA: SNICIT_ELL_WAVE(EPI_DELTA); break;
    case EPI_DELTA_LOSSY: SNICIT_ELL_WAVE(EPI_DELTA_LOSSY); break;
    case EPI_LINEAR: SNICIT_ELL_WAVE(EPI_LINEAR); break;
    default: throw Error(SNICIT_E_INVALID, "bad epilogue");
  }
#undef SNICIT_ELL_WAVE
  check_launch();
}

// 256-column tiles (four columns per lane) once the plan holds at least 4096 columns,
// 64-column tiles below that
void launch_ell_wave(const EllLaunch& a, hipStream_t s) {
  if (a.max_cols >= 4096) launch_ell_wave_w<4>(a, s);
  else launch_ell_wave_w<1>(a, s);
}

// Live column count from which a dense-layer wavefront takes 64 rows x 128 columns (four
// accumulators per input step, one wavefront per SIMD) rather than 32 x 64 (one accumulator,
// four wavefronts per SIMD): below it the 32 x 64 tiles keep more of the CUs busy.
constexpr int64_t kDenseWideCols = 2304;

template <int RB, int CB>
void launch_dense_mfma_t(const EllLaunch& a, int64_t lo, int64_t hi, hipStream_t s) {
  const int64_t row_tiles = a.wt_ld / (32 * RB);
  const int64_t cols = std::min<int64_t>(a.max_cols, hi);
  const int64_t tiles = (cols + 64 * CB - 1) / (64 * CB);
  // four wavefronts per workgroup, or a thread per (row, column) pair when the count may be
  // narrow; at most eight resident workgroups per CU
  const int64_t pairs = lo <= a.shape.narrow ? a.rows * std::min<int64_t>(cols, a.shape.narrow) : 0;
  const int64_t most = std::max<int64_t>({1, (row_tiles * tiles + 3) / 4, (pairs + 255) / 256});
  const unsigned grid = (unsigned)std::min<int64_t>(most, (int64_t)8 * std::max(device_cus(), 1));
#define SNICIT_DENSE(E)                                                                        \
  hipLaunchKernelGGL((dense_mfma_kernel<E, RB, CB>), dim3(grid), dim3(256), 0, s, a.wt,         \
                     a.wt_ld, a.vals, a.K, a.rows, a.bias_vec, a.bias_s, a.clip, a.yin, a.ldin,  \
                     a.yout, a.ldout, a.count, a.shape.narrow, lo, hi)
  if (a.epilogue == EPI_PLAIN) SNICIT_DENSE(EPI_PLAIN);
  else SNICIT_DENSE(EPI_LINEAR);
#undef SNICIT_DENSE
  check_launch();
}

// 32 x 64 tiles, and from kDenseWideCols live columns 64 x 128 tiles (both launched when the
// capacity reaches it and the padded rows are whole 64-row tiles; the (row, column) chains
// stay with the 32 x 64 form)
void launch_dense_mfma(const EllLaunch& a, hipStream_t s) {
  constexpr int64_t kAll = INT64_MAX;
  if (a.max_cols >= kDenseWideCols && a.shape.narrow < kDenseWideCols && a.wt_ld % 64 == 0) {
    launch_dense_mfma_t<1, 1>(a, 0, kDenseWideCols, s);
    launch_dense_mfma_t<2, 2>(a, kDenseWideCols, kAll, s);
  } else {
    launch_dense_mfma_t<1, 1>(a, 0, kAll, s);
  }
}

void launch_ell_spmm(const EllLaunch& a, hipStream_t s) {
  SNICIT_REQUIRE(a.K % 16 == 0 && a.K > 0, "ELL rows must be padded to a multiple of 16");
  SNICIT_REQUIRE(a.in_rows > 0, "ELL launch needs the input row count");
  if (a.shape.ell_chunked == 2 && a.cols && a.rng && a.ent) {
    launch_ell_wave(a, s);
    return;
  }
  // dense layers without per-column flags, gathers or centroid epilogues: matrix cores
  if (a.shape.ell_chunked == 2 && !a.cols && a.wt && !a.map && !a.flags &&
      (a.epilogue == EPI_PLAIN || a.epilogue == EPI_LINEAR) && a.wt_ld % 32 == 0 &&
      a.wt_ld >= a.rows && a.in_rows == a.K) {
    launch_dense_mfma(a, s);
    return;
  }
  const int64_t lds = std::min<int64_t>((int64_t)a.shape.ell_lds_kb * 1024, kLdsPerBlock);
  if (a.rows >= 128 && a.shape.ell_chunked) {
    // wide tiles, input rows staged in chunks that fit the budget (multiple of 16 rows)
    const int64_t rc = std::min<int64_t>(a.in_rows, lds / ((64 + 4) * 4)) / 16 * 16;
    if (rc >= 16) {
      launch_ell_chunk(a, (int32_t)std::min<int64_t>(rc, a.in_rows), s);
      return;
    }
  }
  switch (ell_tile_cols(a.in_rows, lds)) {
    case 32: launch_ell_lds_form<32>(a, s); return;
    case 16: launch_ell_lds_form<16>(a, s); return;
    case 8: launch_ell_lds_form<8>(a, s); return;
    case 4: launch_ell_lds_form<4>(a, s); return;
    default: break;
  }
  const int64_t chunks = (a.rows + 31) / 32;
  const int64_t tiles = (a.max_cols + kBlock - 1) / kBlock;
  int64_t gy = std::max<int64_t>(1, a.shape.grid_target / std::max<int64_t>(chunks, 1));
  gy = std::min<int64_t>(gy, std::max<int64_t>(tiles, 1));
  gy = std::min<int64_t>(gy, 65535);
  dim3 grid((unsigned)chunks, (unsigned)gy);
#define SNICIT_ELL(E)                                                                        \
  hipLaunchKernelGGL(ell_spmm_kernel<E>, grid, dim3(kBlock), 0, s, a.cols, a.vals, a.K, a.rows, \
                     a.bias_vec, a.bias_s, a.clip, a.yin, a.ldin, a.yout, a.ldout, a.count,   \
                     a.map, a.flags, a.zc, a.kcap, a.mcent, a.eps_post, a.shape.narrow)
  switch (a.epilogue) {
    case EPI_PLAIN: SNICIT_ELL(EPI_PLAIN); break;
    case EPI_CENTROID: SNICIT_ELL(EPI_CENTROID); break;
    case EPI_DELTA: SNICIT_ELL(EPI_DELTA); break;
    case EPI_DELTA_LOSSY: SNICIT_ELL(EPI_DELTA_LOSSY); break;
    case EPI_LINEAR: SNICIT_ELL(EPI_LINEAR); break;
    default: throw Error(SNICIT_E_INVALID, "bad epilogue");
  }
#undef SNICIT_ELL
  check_launch();
}

void launch_argmax(const float* logits, int64_t ld, int32_t rows, const int32_t* count,
                   int64_t max_cols, int32_t* cat, hipStream_t s) {
  hipLaunchKernelGGL(argmax_kernel, dim3(grid1(max_cols, 256, 1 << 30)), dim3(256), 0, s, logits,
                     ld, rows, count, cat);
  check_launch();
}

void launch_lossy_pool(const float* Y, int64_t ld, const Segs& sg, int32_t pool, int32_t N,
                       float* poolrows, int32_t* poolcol, hipStream_t s) {
  SNICIT_REQUIRE(pool >= 1 && pool <= kMaxCentroidsLossy, "lossy pool must be in [1, 128]");
  hipLaunchKernelGGL(lossy_pool_kernel, dim3((unsigned)pool, (unsigned)sg.nseg), dim3(256), 0, s,
                     Y, ld, sg.seglo, pool, N, poolrows, poolcol);
  check_launch();
}

void launch_lossy_select(const float* poolrows, const int32_t* poolcol, const Segs& sg,
                         int32_t pool, int32_t N, float eps, uint64_t* close, int32_t* cent_col,
                         hipStream_t s) {
  SNICIT_REQUIRE(pool >= 1 && pool <= sg.stride && pool <= kMaxCentroidsLossy,
                 "pool larger than the centroid stride");
  launch_fill_i32(sg.kcols, 0, 1, s);
  hipLaunchKernelGGL(lossy_close_kernel, dim3((unsigned)pool, (unsigned)sg.nseg), dim3(1024), 0,
                     s, poolrows, sg.seglo, pool, N, tol_cut(eps), (double)N * tol_decimal(eps),
                     close);
  check_launch();
  hipLaunchKernelGGL(lossy_prune_kernel, dim3((unsigned)sg.nseg), dim3(kWave), 0, s, close,
                     poolcol, sg.seglo, pool, sg.stride, cent_col, sg.kseg, sg.kcols);
  check_launch();
}

void launch_lossy_assign(const float* Y, int64_t ld, const Segs& sg, int64_t max_cols,
                         const float* C, int32_t kcap, const int32_t* cent_col, int32_t N,
                         float eps, int32_t* dist, int32_t* mcar, int32_t* centmark, int32_t* rec,
                         hipStream_t s) {
  SNICIT_REQUIRE(sg.stride % 8 == 0 && sg.stride <= kMaxCentroidsLossy &&
                     kcap >= sg.stride * sg.nseg,
                 "bad centroid layout");
  // (64-column tiles, blocks of eight centroids; a block past a segment's k does nothing)
  dim3 grid(grid1(max_cols, kWave, 1 << 30), (unsigned)(sg.stride / 8));
  hipLaunchKernelGGL(lossy_dist_kernel, grid, dim3(kDistWaves * kWave), 0, s, Y, ld, sg.seglo,
                     sg.nseg, C, kcap, sg.stride, sg.kseg, N, tol_cut(eps), dist);
  check_launch();
  hipLaunchKernelGGL(lossy_final_kernel, dim3(grid1(max_cols, 256, 1 << 30)), dim3(256), 0, s,
                     sg.seglo, sg.nseg, sg.kseg, sg.stride, cent_col, dist, ld, mcar, centmark,
                     rec);
  check_launch();
}

void launch_lossy_delta(const float* Y, int64_t ld, const float* C, int32_t kcap,
                        const int32_t* mcar, const int32_t* centmark, const int32_t* B,
                        int64_t max_cols, int32_t N, float eps, bool positive_only, float* D,
                        int64_t ldD, int32_t* flags, hipStream_t s) {
  dim3 grid(grid1(max_cols, 256, 1 << 30), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(lossy_delta_kernel, grid, dim3(256), 0, s, Y, ld, C, kcap, mcar, centmark, B,
                     N, tol_cut(eps), positive_only ? 1 : 0, D, ldD, flags);
  check_launch();
}

void launch_compact(int32_t* flags, const int32_t* count_in, int32_t* count_out, int32_t* map_out,
                    const int32_t* compose, const int32_t* c0_in, int32_t* c0_out,
                    const int32_t* c1_in, int32_t* c1_out, int64_t max_cols, int32_t* tile_cnt,
                    hipStream_t s, int32_t slack, int32_t* live_out) {
  SNICIT_REQUIRE(slack >= 0 && slack < 1024, "compaction slack out of range");
  SNICIT_REQUIRE(max_cols >= 0 && max_cols < (int64_t(1) << 31), "compaction size out of range");
  // tiles of kCompactTile columns when the count may exceed one workgroup's walk; without
  // the tile-count scratch (lists that stay short) one workgroup walks any count
  const unsigned tiles = max_cols > kCompactOneBlock && tile_cnt
                             ? (unsigned)((max_cols + kCompactTile - 1) / kCompactTile)
                             : 1u;
  if (tiles > 1) {
    hipLaunchKernelGGL(compact_count_kernel, dim3(tiles), dim3(1024), 0, s, flags, count_in,
                       tile_cnt);
    check_launch();
  }
  hipLaunchKernelGGL(compact_kernel, dim3(tiles), dim3(1024), 0, s, flags, count_in, count_out,
                     map_out, compose, c0_in, c0_out, c1_in, c1_out, tile_cnt, slack, live_out);
  check_launch();
}

void launch_ring_count(int32_t* ring, int64_t ring_ld, int32_t j0, int32_t nslots,
                       const int32_t* count, int32_t* nerow, hipStream_t s) {
  SNICIT_REQUIRE(nslots >= 1 && nslots <= kPostRing, "ring range");
  hipLaunchKernelGGL(ring_count_kernel, dim3((unsigned)nslots), dim3(1024), 0, s, ring, ring_ld,
                     j0, nslots, count, nerow);
  check_launch();
}
void launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(fill_i32_kernel, dim3(grid1(n, 256, 4096)), dim3(256), 0, s, p, v, n);
  check_launch();
}
void launch_iota_i32(int32_t* p, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(iota_i32_kernel, dim3(grid1(n, 256, 4096)), dim3(256), 0, s, p, n);
  check_launch();
}
void launch_fill_f32(float* p, float v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(fill_f32_kernel, dim3(grid1(n, 256, 65535)), dim3(256), 0, s, p, v, n);
  check_launch();
}

void launch_count_nonzero(const float* y, int64_t rows, int64_t ld, int64_t B,
                          unsigned long long* out, hipStream_t s) {
  const int64_t waves = rows * ((B + 63) / 64);
  hipLaunchKernelGGL(count_nonzero_kernel, dim3(grid1(waves * 64, 256, 8192)), dim3(256), 0, s,
                     y, rows, ld, B, out);
  check_launch();
}

void launch_transpose_in(const float* src, int64_t B, int32_t N, float* dst, int64_t ld,
                         int32_t* nonfinite, hipStream_t s) {
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)((B + 63) / 64));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, src, B, (int64_t)N, (int64_t)N,
                     dst, ld, nonfinite);
  check_launch();
}
void launch_transpose_out(const float* src, int64_t ld, int64_t B, int32_t N, float* dst,
                          hipStream_t s) {
  dim3 grid((unsigned)((B + 63) / 64), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, src, (int64_t)N, B, ld, dst,
                     (int64_t)N, (int32_t*)nullptr);
  check_launch();
}
void launch_scatter_cols(const float* src, int64_t ldsrc, const int32_t* map,
                         const int32_t* dstcol, const int32_t* count, int64_t max_cols,
                         int32_t N, float* dst, int64_t lddst, hipStream_t s) {
  dim3 grid(grid1(max_cols, 256, 1 << 30), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(scatter_cols_kernel, grid, dim3(256), 0, s, src, ldsrc, map, dstcol, count,
                     N, dst, lddst);
  check_launch();
}

void launch_images(const ImageGeom& g, const int32_t* d_params, const int32_t* d_blobs,
                   const int32_t* d_proto, int64_t B, int32_t N, float* Y, int64_t ld,
                   hipStream_t s) {
  dim3 grid(grid1(B, 256, 1 << 30), (unsigned)((N + 255) / 256));
  hipLaunchKernelGGL(images_kernel, grid, dim3(256), 0, s, d_params, d_blobs, d_proto, g.W, g.H,
                     B, N, Y, ld);
  check_launch();
}

void launch_row_any(const float* Y, int64_t ld, const int32_t* count, int64_t max_cols,
                    int32_t N, int32_t* flags, hipStream_t s) {
  dim3 grid(grid1(max_cols, 256, 1 << 30), (unsigned)((N + 255) / 256));
  hipLaunchKernelGGL(row_any_kernel, grid, dim3(256), 0, s, Y, ld, count, N, flags);
  check_launch();
}
void launch_cat_from_list(const int32_t* orig, const int32_t* count, int64_t max_cols,
                          int32_t* cat, hipStream_t s) {
  hipLaunchKernelGGL(cat_from_list_kernel, dim3(grid1(max_cols, 256, 4096)), dim3(256), 0, s,
                     orig, count, cat);
  check_launch();
}
void launch_cat_from_flags(int32_t* flags, const int32_t* count, int64_t max_cols, int32_t* cat,
                           hipStream_t s) {
  hipLaunchKernelGGL(cat_from_flags_kernel, dim3(grid1(max_cols, 256, 4096)), dim3(256), 0, s,
                     flags, count, cat);
  check_launch();
}

void launch_segment_bounds(const int32_t* orig, const int32_t* count, const int32_t* segstart,
                           const Segs& sg, hipStream_t s) {
  hipLaunchKernelGGL(segment_bounds_kernel, dim3(1), dim3(256), 0, s, orig, count, segstart,
                     sg.nseg, sg.seglo);
  check_launch();
}
void launch_seg_sums(const float* Y, int64_t ld, const int32_t* map, const Segs& sg, int32_t pool,
                     int32_t segments, int32_t N, float* F, hipStream_t s) {
  SNICIT_REQUIRE(pool <= 32 && segments <= 32, "pool and segments must be <= 32");
  hipLaunchKernelGGL(seg_sums_kernel, dim3((unsigned)pool, (unsigned)sg.nseg), dim3(1024), 0, s, Y,
                     ld, map, sg.seglo, pool, segments, N, F);
  check_launch();
}
void launch_select_exact(const float* F, const Segs& sg, int32_t pool, int32_t segments, float eps,
                         int32_t* cent_slot, hipStream_t s) {
  SNICIT_REQUIRE(pool <= sg.stride && sg.stride <= 64, "bad exact centroid layout");
  launch_fill_i32(sg.kcols, 0, 1, s);
  hipLaunchKernelGGL(select_exact_kernel, dim3((unsigned)sg.nseg), dim3(64), 0, s, F, sg.seglo,
                     sg.nseg, pool, segments, tol_cut(eps), (double)segments * tol_decimal(eps),
                     sg.stride, cent_slot, sg.kseg, sg.kcols);
  check_launch();
}
void launch_gather_centroids(const float* Y, int64_t ld, const int32_t* map,
                             const int32_t* cent_slot, const Segs& sg, int32_t N, float* C,
                             int32_t kcap, hipStream_t s) {
  SNICIT_REQUIRE(kcap >= sg.stride * sg.nseg, "bad centroid layout");
  hipLaunchKernelGGL(gather_centroids_kernel, dim3(grid1((int64_t)N * kcap, 256, 1 << 30)),
                     dim3(256), 0, s, Y, ld, map, cent_slot, sg.kseg, sg.stride, N, C, kcap);
  check_launch();
}
void launch_assign_exact(const float* Y, int64_t ld, const int32_t* map, const Segs& sg,
                         int64_t max_cols, const float* C, int32_t kcap, const int32_t* cent_slot,
                         int32_t N, int32_t* dist, int32_t* flags, int32_t* mcar,
                         const int32_t* orig, int32_t* rec, hipStream_t s) {
  SNICIT_REQUIRE(sg.stride == kMaxCentroidsExact && kcap >= sg.stride * sg.nseg,
                 "assign expects 32 centroid slots per segment");
  const int nchunk = std::max(1, std::min(64, N / 128));
  dim3 grid(grid1(max_cols, 256, 1 << 30), (unsigned)nchunk);
  hipLaunchKernelGGL(assign_dist_kernel, grid, dim3(256), 0, s, Y, ld, map, sg.seglo, sg.nseg, C,
                     kcap, sg.stride, sg.kseg, N, nchunk, dist);
  check_launch();
  hipLaunchKernelGGL(assign_final_kernel, dim3(grid1(max_cols, 256, 1 << 30)), dim3(256), 0, s,
                     sg.seglo, sg.nseg, sg.kseg, sg.stride, cent_slot, dist, flags, mcar, orig,
                     rec);
  check_launch();
}
void launch_make_delta(const float* Y, int64_t ld, const int32_t* map, const float* C,
                       int32_t kcap, const int32_t* m, const int32_t* count, int64_t max_cols,
                       int32_t N, float* D, int64_t ldD, bool compact, hipStream_t s) {
  dim3 grid(grid1(max_cols, 256, 1 << 30), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(make_delta_kernel, grid, dim3(256), 0, s, Y, ld, map, C, kcap, m, count, N,
                     D, ldD, compact ? 1 : 0);
  check_launch();
}
void launch_centroid_cats(const float* C, int32_t kcap, const int32_t* k, int32_t N,
                          int32_t* catc, hipStream_t s) {
  hipLaunchKernelGGL(centroid_cats_kernel, dim3((unsigned)kcap), dim3(256), 0, s, C, kcap, k, N,
                     catc);
  check_launch();
}
void launch_recover_base(const int32_t* rec, const int32_t* B, int64_t max_cols,
                         const int32_t* catc, int32_t* cat, hipStream_t s) {
  hipLaunchKernelGGL(recover_base_kernel, dim3(grid1(max_cols, 256, 4096)), dim3(256), 0, s, rec,
                     B, catc, cat);
  check_launch();
}
void launch_recover_rows(const float* D, int64_t ldD, const int32_t* ldD_cnt, const int32_t* map,
                         const int32_t* m, const int32_t* orig, const int32_t* count,
                         int64_t max_cols, const float* C, int32_t kcap, int32_t N, int32_t* act,
                         int32_t* cat, hipStream_t s) {
  dim3 grid(grid1(max_cols, 256, 1 << 30), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(recover_rows_kernel, grid, dim3(256), 0, s, D, ldD, ldD_cnt, map, m, count,
                     C, kcap, N, act);
  check_launch();
  hipLaunchKernelGGL(recover_write_kernel, dim3(grid1(max_cols, 256, 4096)), dim3(256), 0, s,
                     orig, count, act, cat);
  check_launch();
}
void launch_expand_centroids(const int32_t* rec, const int32_t* B, int64_t max_cols,
                             const float* C, int32_t kcap, int32_t N, float* Y, int64_t ld,
                             hipStream_t s) {
  dim3 grid(grid1(max_cols, 256, 1 << 30), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(expand_centroids_kernel, grid, dim3(256), 0, s, rec, B, C, kcap, N, Y, ld);
  check_launch();
}
void launch_scatter_delta(const float* D, int64_t ldD, const int32_t* ldD_cnt, const int32_t* map,
                          const int32_t* m, const int32_t* orig, const int32_t* count,
                          int64_t max_cols, const float* C, int32_t kcap, int32_t N, float* Y,
                          int64_t ld, hipStream_t s) {
  dim3 grid(grid1(max_cols, 256, 1 << 30), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(scatter_delta_kernel, grid, dim3(256), 0, s, D, ldD, ldD_cnt, map, m, orig,
                     count, C, kcap, N, Y, ld);
  check_launch();
}

}  // namespace snicit
