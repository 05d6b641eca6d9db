/*
 * oracle.c -- CPU restatement of the reference's SNICIT layer loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load this library, and only as the checker or as the timed CPU
 * baseline; the product (snicit_amd/libsnicit.so) never calls it.
 *
 * Layout: sample-major, y[b * N + n].  Nets are per-row ELL: row o of layer l has K
 * entries cols[(l*N+o)*K + r] (input neuron) with value vals[...].
 *
 * Arithmetic contract shared with the GPU build (so FP32 results compare bit for bit):
 *   z[o] = fmaf chain over the row's entries in ascending input neuron (stable for ties),
 *          starting from +0: acc = fmaf(w, y, acc);
 *   plain:     y' = min(max(z + b, 0), clip)
 *   centroid:  zc = z + b,  y' = min(max(zc, 0), clip)
 *   delta:     y' = f(zc[m] + z) - f(zc[m])
 * The reference accumulates with shared-memory float atomics in arbitrary order
 * (src/SDGC/BF/kernel.hpp:73) or sequential FMAs in its own layout order
 * (src/SDGC/SNICIT/kernel.hpp:94-98); any fixed order is a faithful restatement, and the
 * ascending order is the one both sides here use.
 *
 * Tolerances: the reference writes them as double literals (0.03, 0.04, 0.05) and compares
 * float magnitudes against them in double (src/SDGC/SNICIT/kernel.hpp:178,184;
 * src/beyond/SNICIT/kernel.hpp:31,38,90,161).  They reach this library as floats; the value
 * compared against is the shortest decimal that rounds to the given float (0.05f -> 0.05),
 * so |x| = 0.05f counts as above 0.05, as in the reference.
 *
 * Parity pins (tests/test_cpu.py, tests/test_cpu_cluster.py, tests/test_cpu_beyond.py):
 * closed forms (layer 0 on binary inputs, dead-bias nets, a permutation net, dense brute
 * force), invariants (SNICIT categories == plain categories for any centroid choice;
 * yhat_i + yhat_m(i) == y_i), numpy recomputation of the clustering stages (segment sums,
 * the Alg. 1 survivor characterisation, first-argmin assignment), hand-built clustering
 * cases with known answers, FP64 agreement, and the fixtures under tests/golden/.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SNICIT_BIAS_SCALAR 0
#define SNICIT_BIAS_PER_LAYER 1
#define SNICIT_BIAS_PER_NEURON 2

/* f(z) = min(max(z,0),clip): src/SDGC/XY/kernel.hpp:5-7 (clip 32),
 * src/beyond/general.hpp:70 (clip 1). */
static inline float act(float z, float clip) { return fminf(fmaxf(z, 0.f), clip); }

/* Decimal value of a tolerance given as a float: the shortest decimal that rounds to it. */
static double decimal_of(float e) {
  char buf[32];
  for (int p = 1; p <= 9; ++p) {
    snprintf(buf, sizeof buf, "%.*g", p, (double)e);
    if (strtof(buf, NULL) == e) return strtod(buf, NULL);
  }
  return (double)e;
}
static inline double actd(double z, double clip) { return fmin(fmax(z, 0.0), clip); }

static float bias_of(int bias_kind, const float* bias, int N, int l, int o) {
  if (bias_kind == SNICIT_BIAS_SCALAR) return bias[0];
  if (bias_kind == SNICIT_BIAS_PER_LAYER) return bias[l];
  return bias[(size_t)l * N + o];
}

static int all_bias_nonpositive(int bias_kind, const float* bias, int N, int L) {
  size_t n = bias_kind == SNICIT_BIAS_SCALAR ? 1 : bias_kind == SNICIT_BIAS_PER_LAYER ? (size_t)L : (size_t)L * N;
  for (size_t i = 0; i < n; ++i)
    if (!(bias[i] <= 0.f)) return 0;
  return 1;
}

/* Row entries sorted by input neuron (stable): idx[(l*N+o)*K + r] = original slot. */
typedef struct {
  int N, L, K;
  int32_t* scol; /* sorted columns */
  float* sval;   /* matching values */
} sorted_net;

static int cmp_pair(const void* a, const void* b) {
  const int64_t* x = (const int64_t*)a;
  const int64_t* y = (const int64_t*)b;
  return (*x > *y) - (*x < *y);
}

static int sort_net(int N, int L, int K, const int32_t* cols, const float* vals, sorted_net* s) {
  s->N = N;
  s->L = L;
  s->K = K;
  size_t tot = (size_t)L * N * K;
  s->scol = (int32_t*)malloc(tot * sizeof(int32_t));
  s->sval = (float*)malloc(tot * sizeof(float));
  if (!s->scol || !s->sval) return -1;
#pragma omp parallel
  {
    int64_t* key = (int64_t*)malloc((size_t)K * sizeof(int64_t));
#pragma omp for schedule(static)
    for (int64_t row = 0; row < (int64_t)L * N; ++row) {
      const int32_t* c = cols + (size_t)row * K;
      /* key = column * K + slot: ascending column, stable in slot */
      for (int r = 0; r < K; ++r) key[r] = (int64_t)c[r] * K + r;
      qsort(key, K, sizeof(int64_t), cmp_pair);
      for (int r = 0; r < K; ++r) {
        int slot = (int)(key[r] % K);
        s->scol[(size_t)row * K + r] = c[slot];
        s->sval[(size_t)row * K + r] = vals[(size_t)row * K + slot];
      }
    }
    free(key);
  }
  return 0;
}

static void free_net(sorted_net* s) {
  free(s->scol);
  free(s->sval);
}

/* z = W_l y (no bias), fixed ascending-column fmaf chain. */
static inline float row_dot(const sorted_net* s, int l, int o, const float* y) {
  const size_t base = ((size_t)l * s->N + o) * s->K;
  const int32_t* c = s->scol + base;
  const float* w = s->sval + base;
  float acc = 0.f;
  for (int r = 0; r < s->K; ++r) acc = fmaf(w[r], y[c[r]], acc);
  return acc;
}

static int row_nonzero(const float* y, int N) {
  for (int n = 0; n < N; ++n)
    if (y[n] != 0.f) return 1;
  return 0;
}

/* ---- plain forward (BF, src/SDGC/BF/kernel.hpp:26-89; BF.hpp:157-220) -------------
 * Y_{l+1}[b,o] = f(W_l Y_l[b] + b_l).  With every bias <= 0 an all-zero row stays zero, so
 * those rows are skipped (BF's rowsY, BF.hpp:222-229).  category(b) = any(Y_L[b] > 0)
 * (identify, src/SDGC/utility.hpp:194-212: sum > 0 with y >= 0).
 * live_per_layer[l] = rows with a nonzero after layer l (nullable). */
int oracle_plain_f32_ex(int N, int L, int K, const int32_t* cols, const float* vals,
                        int bias_kind, const float* bias, float clip, int64_t B, const float* y0,
                        float* yL, int32_t* cat, int64_t* live_per_layer, int nthreads,
                        double* forward_seconds);

int oracle_plain_f32(int N, int L, int K, const int32_t* cols, const float* vals, int bias_kind,
                     const float* bias, float clip, int64_t B, const float* y0, float* yL,
                     int32_t* cat, int64_t* live_per_layer, int nthreads) {
  return oracle_plain_f32_ex(N, L, K, cols, vals, bias_kind, bias, clip, B, y0, yL, cat,
                             live_per_layer, nthreads, NULL);
}

/* As oracle_plain_f32; forward_seconds (nullable) gets the wall time of the layer loop
 * alone (the one-off sort of the rows is excluded). */
int oracle_plain_f32_ex(int N, int L, int K, const int32_t* cols, const float* vals,
                        int bias_kind, const float* bias, float clip, int64_t B, const float* y0,
                        float* yL, int32_t* cat, int64_t* live_per_layer, int nthreads,
                        double* forward_seconds) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  sorted_net s;
  if (sort_net(N, L, K, cols, vals, &s)) return -1;
  const int skip = all_bias_nonpositive(bias_kind, bias, N, L);
  float* a = (float*)malloc((size_t)B * N * sizeof(float));
  float* b = (float*)malloc((size_t)B * N * sizeof(float));
  unsigned char* live = (unsigned char*)malloc((size_t)B);
  if (!a || !b || !live) return -1;
  memcpy(a, y0, (size_t)B * N * sizeof(float));
  for (int64_t i = 0; i < B; ++i) live[i] = 1;
  const double t0 = omp_get_wtime();
  for (int l = 0; l < L; ++l) {
    int64_t nlive = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : nlive)
    for (int64_t i = 0; i < B; ++i) {
      float* out = b + (size_t)i * N;
      if (skip && !live[i]) {
        memset(out, 0, (size_t)N * sizeof(float));
        continue;
      }
      const float* in = a + (size_t)i * N;
      int any = 0;
      for (int o = 0; o < N; ++o) {
        float y = act(row_dot(&s, l, o, in) + bias_of(bias_kind, bias, N, l, o), clip);
        out[o] = y;
        any |= y > 0.f;
      }
      live[i] = (unsigned char)any;
      nlive += any;
    }
    if (live_per_layer) live_per_layer[l] = nlive;
    float* t = a;
    a = b;
    b = t;
  }
  if (forward_seconds) *forward_seconds = omp_get_wtime() - t0;
  for (int64_t i = 0; i < B; ++i) {
    const float* y = a + (size_t)i * N;
    int any = 0;
    for (int n = 0; n < N && !any; ++n) any = y[n] > 0.f;
    if (cat) cat[i] = any;
  }
  if (yL) memcpy(yL, a, (size_t)B * N * sizeof(float));
  free(a);
  free(b);
  free(live);
  free_net(&s);
  return 0;
}

/* ---- FP64 forward: the same net in double (accumulation order is immaterial at this
 * precision for the <= 1e-4 comparisons it serves). */
int oracle_plain_f64(int N, int L, int K, const int32_t* cols, const float* vals, int bias_kind,
                     const float* bias, float clip, int64_t B, const float* y0, double* yL,
                     int32_t* cat, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double* a = (double*)malloc((size_t)B * N * sizeof(double));
  double* b = (double*)malloc((size_t)B * N * sizeof(double));
  if (!a || !b) return -1;
  for (size_t i = 0; i < (size_t)B * N; ++i) a[i] = y0[i];
  for (int l = 0; l < L; ++l) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < B; ++i) {
      const double* in = a + (size_t)i * N;
      double* out = b + (size_t)i * N;
      for (int o = 0; o < N; ++o) {
        const size_t base = ((size_t)l * N + o) * K;
        double acc = 0.0;
        for (int r = 0; r < K; ++r) acc += (double)vals[base + r] * in[cols[base + r]];
        out[o] = actd(acc + (double)bias_of(bias_kind, bias, N, l, o), (double)clip);
      }
    }
    double* t = a;
    a = b;
    b = t;
  }
  for (int64_t i = 0; i < B; ++i) {
    int any = 0;
    for (int n = 0; n < N && !any; ++n) any = a[(size_t)i * N + n] > 0.0;
    if (cat) cat[i] = any;
  }
  if (yL) memcpy(yL, a, (size_t)B * N * sizeof(double));
  free(a);
  free(b);
  return 0;
}

/* One layer in FP64 from an FP32 input (layer-local error checks). */
int oracle_layer_f64(int N, int K, const int32_t* cols_l, const float* vals_l, int bias_kind,
                     const float* bias_l, float clip, int64_t B, const float* yin, double* yout) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < B; ++i) {
    const float* in = yin + (size_t)i * N;
    for (int o = 0; o < N; ++o) {
      const size_t base = (size_t)o * K;
      double acc = 0.0;
      for (int r = 0; r < K; ++r) acc += (double)vals_l[base + r] * (double)in[cols_l[base + r]];
      double b = bias_kind == SNICIT_BIAS_PER_NEURON ? bias_l[o] : bias_l[0];
      yout[(size_t)i * N + o] = actd(acc + b, (double)clip);
    }
  }
  return 0;
}

/* ---- exact SNICIT (src/SDGC/SNICIT/SNICIT.hpp:205-692, kernel.hpp:8-218) ---------------
 * Pre-threshold: plain layers 0..t-1 with dead-row skipping.
 * At t, over the live rows B' (ascending sample id; SNICIT.hpp:569-588 compaction):
 *   pool  = first min(sample_size, B') live rows (reduction, kernel.hpp:138-153; intended
 *           segments of N/segments neurons each, summed in ascending neuron order)
 *   Alg. 1 (centroid_sel, kernel.hpp:156-194): in pool order, a present base c removes
 *           every other present row y with #{g : |F_y[g] - F_c[g]| > eps} < segments*eps
 *   m(i)  = first argmin_c #{n : Y_i[n] != Y_c[n]} (coarse_cluster, kernel.hpp:27-49);
 *           centroids map to themselves
 *   delta = Y_i - Y_m(i) (kernel.hpp:50-58, with the stored m, quirk Q1 not reproduced)
 *   ne(i) = non-centroid with a nonzero distance (kernel.hpp:59-65)
 * Post-threshold (post_spMM + post_minus, kernel.hpp:71-136) for l = t..L-1:
 *   Zc = W yhat_c + b, yhat_c' = f(Zc); for ne rows: yhat_i' = f(Zc[m] + W yhat_i) - f(Zc[m]),
 *   ne(i) = any(yhat_i' != 0) (recompacted every layer; the reference every 200).
 * Recovery (recover, kernel.hpp:197-218): cat(i) = any(yhat_i + yhat_m(i) > 0); rows dead
 * before t are 0.  yL (nullable) gets yhat_i + yhat_m(i) (or yhat_m(i) for empty rows).
 * Trace outputs (all nullable): pool_ids[sample_size], F[sample_size*segments], cent_ids[32],
 * assign[B] (centroid index, -1 dead before t), live_per_layer[L] (layers < t),
 * nerow_per_layer[L] (after each post layer), info[4] = {B', npool, k, nerow at t}.
 * seg_layout 0: segment g = neurons [g*N/S, (g+1)*N/S) (paper 3.2.1).  seg_layout 1
 * reproduces the reference's indexing (quirk Q2, kernel.hpp:147-148, launched with
 * blockDim (1024/S, S) at SNICIT.hpp:434-437): segment g = neurons [g*1024/S, g*1024/S + N/S),
 * which overlap for N > 1024; it needs 1024 % S == 0 and (S-1)*1024/S + N/S <= N. */
int oracle_snicit_exact_ex(int N, int L, int K, const int32_t* cols, const float* vals,
                           int bias_kind, const float* bias, float clip, int64_t B, const float* y0,
                           int t, int sample_size, int segments, float eps, int seg_layout,
                           int32_t* cat, float* yL, int32_t* pool_ids, float* F_out,
                           int32_t* cent_ids, int32_t* assign, int64_t* live_per_layer,
                           int64_t* nerow_per_layer, int64_t* info, int nthreads);

int oracle_snicit_exact(int N, int L, int K, const int32_t* cols, const float* vals,
                        int bias_kind, const float* bias, float clip, int64_t B, const float* y0,
                        int t, int sample_size, int segments, float eps, int32_t* cat, float* yL,
                        int32_t* pool_ids, float* F_out, int32_t* cent_ids, int32_t* assign,
                        int64_t* live_per_layer, int64_t* nerow_per_layer, int64_t* info,
                        int nthreads) {
  return oracle_snicit_exact_ex(N, L, K, cols, vals, bias_kind, bias, clip, B, y0, t, sample_size,
                                segments, eps, 0, cat, yL, pool_ids, F_out, cent_ids, assign,
                                live_per_layer, nerow_per_layer, info, nthreads);
}

int oracle_snicit_exact_ex(int N, int L, int K, const int32_t* cols, const float* vals,
                           int bias_kind, const float* bias, float clip, int64_t B, const float* y0,
                           int t, int sample_size, int segments, float eps, int seg_layout,
                           int32_t* cat, float* yL, int32_t* pool_ids, float* F_out,
                           int32_t* cent_ids, int32_t* assign, int64_t* live_per_layer,
                           int64_t* nerow_per_layer, int64_t* info, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (t < 0 || t >= L || sample_size < 1 || sample_size > 64 || segments < 1 || N % segments) return -2;
  if (seg_layout != 0 && (seg_layout != 1 || 1024 % segments ||
                          (int64_t)(segments - 1) * (1024 / segments) + N / segments > N))
    return -2;
  sorted_net s;
  if (sort_net(N, L, K, cols, vals, &s)) return -1;
  const int skip = all_bias_nonpositive(bias_kind, bias, N, L);
  float* a = (float*)malloc((size_t)B * N * sizeof(float));
  float* b2 = (float*)malloc((size_t)B * N * sizeof(float));
  unsigned char* live = (unsigned char*)malloc((size_t)B + 1);
  if (!a || !b2 || !live) return -1;
  memcpy(a, y0, (size_t)B * N * sizeof(float));
  for (int64_t i = 0; i < B; ++i) live[i] = 1;
  for (int l = 0; l < t; ++l) {
    int64_t nlive = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : nlive)
    for (int64_t i = 0; i < B; ++i) {
      float* out = b2 + (size_t)i * N;
      if (skip && !live[i]) {
        memset(out, 0, (size_t)N * sizeof(float));
        continue;
      }
      const float* in = a + (size_t)i * N;
      int any = 0;
      for (int o = 0; o < N; ++o) {
        float y = act(row_dot(&s, l, o, in) + bias_of(bias_kind, bias, N, l, o), clip);
        out[o] = y;
        any |= y > 0.f;
      }
      live[i] = (unsigned char)any;
      nlive += any;
    }
    if (live_per_layer) live_per_layer[l] = nlive;
    float* tmp = a;
    a = b2;
    b2 = tmp;
  }
  /* live rows at t */
  int64_t* lid = (int64_t*)malloc((size_t)(B + 1) * sizeof(int64_t));
  int64_t Bp = 0;
  for (int64_t i = 0; i < B; ++i) {
    int lv = skip ? row_nonzero(a + (size_t)i * N, N) : 1;
    if (lv) lid[Bp++] = i;
  }
  const int P = (int)(Bp < sample_size ? Bp : sample_size);
  const int seglen = N / segments;
  float* F = (float*)calloc((size_t)(P > 0 ? P : 1) * segments, sizeof(float));
  /* F: segment g split into S2 = min(32, seglen) contiguous sub-chunks, each summed in
   * ascending neuron order, partials added in ascending order (a fixed order for the
   * reference's unordered shared-memory atomics, kernel.hpp:147-150).  The segment starts
   * at g*seglen, or at g*(1024/segments) with the reference's indexing. */
  const int S2 = seglen < 32 ? seglen : 32;
  const int seg_step = seg_layout == 1 ? 1024 / segments : seglen;
  for (int q = 0; q < P; ++q) {
    const float* y = a + (size_t)lid[q] * N;
    for (int g = 0; g < segments; ++g) {
      float sum = 0.f;
      for (int c = 0; c < S2; ++c) {
        float part = 0.f;
        for (int n = c * seglen / S2; n < (c + 1) * seglen / S2; ++n) part += y[g * seg_step + n];
        sum += part;
      }
      F[q * segments + g] = sum;
    }
  }
  const double eps_d = decimal_of(eps);
  int alive[64];
  for (int q = 0; q < 64; ++q) alive[q] = q < P;
  for (int c = 0; c < P; ++c) {
    if (!alive[c]) continue;
    for (int y = 0; y < P; ++y) {
      if (!alive[y] || y == c) continue;
      int cnt = 0;
      for (int g = 0; g < segments; ++g)
        cnt += (double)fabsf(F[y * segments + g] - F[c * segments + g]) > eps_d;
      if ((double)cnt < (double)segments * eps_d) alive[y] = 0;
    }
  }
  int cent[64], k = 0;
  for (int q = 0; q < P; ++q)
    if (alive[q]) cent[k++] = q;
  /* centroid activations (copies: the delta rows refer to them) */
  float* C = (float*)malloc((size_t)(k > 0 ? k : 1) * N * sizeof(float));
  for (int c = 0; c < k; ++c) memcpy(C + (size_t)c * N, a + (size_t)lid[cent[c]] * N, (size_t)N * sizeof(float));
  int32_t* m = (int32_t*)malloc((size_t)(Bp + 1) * sizeof(int32_t));
  unsigned char* ne = (unsigned char*)malloc((size_t)Bp + 1);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t j = 0; j < Bp; ++j) {
    const float* y = a + (size_t)lid[j] * N;
    int best = 0x7fffffff, bestc = 0;
    for (int c = 0; c < k; ++c) {
      const float* yc = C + (size_t)c * N;
      int d = 0;
      for (int n = 0; n < N; ++n) d += y[n] != yc[n];
      if (d < best) {
        best = d;
        bestc = c;
      }
    }
    int is_cent = 0;
    for (int c = 0; c < k; ++c)
      if (cent[c] == j) {
        is_cent = 1;
        bestc = c;
      }
    m[j] = bestc;
    ne[j] = (unsigned char)(!is_cent && best > 0);
  }
  /* delta rows: Yhat_i = Y_i - Y_m(i) for ne rows (kept in place of Y_i), zero otherwise */
  int64_t nerow = 0;
  for (int64_t j = 0; j < Bp; ++j) {
    float* y = a + (size_t)lid[j] * N;
    if (ne[j]) {
      const float* yc = C + (size_t)m[j] * N;
      for (int n = 0; n < N; ++n) y[n] = y[n] - yc[n];
      ++nerow;
    }
  }
  if (info) {
    info[0] = Bp;
    info[1] = P;
    info[2] = k;
    info[3] = nerow;
  }
  if (pool_ids)
    for (int q = 0; q < P; ++q) pool_ids[q] = (int32_t)lid[q];
  if (F_out) memcpy(F_out, F, (size_t)P * segments * sizeof(float));
  if (cent_ids)
    for (int c = 0; c < k; ++c) cent_ids[c] = (int32_t)lid[cent[c]];
  float* Zc = (float*)malloc((size_t)(k > 0 ? k : 1) * N * sizeof(float));
  float* Cn = (float*)malloc((size_t)(k > 0 ? k : 1) * N * sizeof(float));
  float* tmprow = NULL;
  for (int l = t; l < L; ++l) {
    for (int c = 0; c < k; ++c) {
      for (int o = 0; o < N; ++o) {
        float z = row_dot(&s, l, o, C + (size_t)c * N) + bias_of(bias_kind, bias, N, l, o);
        Zc[(size_t)c * N + o] = z;
        Cn[(size_t)c * N + o] = act(z, clip);
      }
    }
    int64_t cnt = 0;
#pragma omp parallel reduction(+ : cnt)
    {
      float* outrow = (float*)malloc((size_t)N * sizeof(float));
#pragma omp for schedule(dynamic, 4)
      for (int64_t j = 0; j < Bp; ++j) {
        if (!ne[j]) continue;
        float* y = a + (size_t)lid[j] * N;
        const float* zc = Zc + (size_t)m[j] * N;
        int any = 0;
        for (int o = 0; o < N; ++o) {
          float z = row_dot(&s, l, o, y);
          float v = act(zc[o] + z, clip) - act(zc[o], clip);
          outrow[o] = v;
          any |= v != 0.f;
        }
        memcpy(y, outrow, (size_t)N * sizeof(float));
        ne[j] = (unsigned char)any;
        cnt += any;
      }
      free(outrow);
    }
    if (nerow_per_layer) nerow_per_layer[l] = cnt;
    float* sw = C;
    C = Cn;
    Cn = sw;
  }
  free(tmprow);
  /* recovery */
  if (assign)
    for (int64_t i = 0; i < B; ++i) assign[i] = -1;
  for (int64_t i = 0; i < B; ++i) {
    if (cat) cat[i] = 0;
    if (yL) memset(yL + (size_t)i * N, 0, (size_t)N * sizeof(float));
  }
  for (int64_t j = 0; j < Bp; ++j) {
    const int64_t i = lid[j];
    const float* yc = C + (size_t)m[j] * N;
    const float* y = a + (size_t)i * N;
    int any = 0;
    for (int n = 0; n < N; ++n) {
      float v = ne[j] ? y[n] + yc[n] : yc[n];
      any |= v > 0.f;
      if (yL) yL[(size_t)i * N + n] = v;
    }
    if (cat) cat[i] = any;
    if (assign) assign[i] = m[j];
  }
  free(Zc);
  free(Cn);
  free(C);
  free(m);
  free(ne);
  free(F);
  free(lid);
  free(a);
  free(b2);
  free(live);
  free_net(&s);
  return 0;
}

/* ---- beyond-style models: dense input layer, hidden layers, dense output + argmax -------
 * Dense input (src/beyond/general.hpp:5-33): h = min(max(W_in x + b_in, 0), in_clip), rows of
 * W_in are [N][Nin], chain over inputs 0..Nin-1.  Dense output (general.hpp:74-102): logits =
 * W_out y + b_out ([Nout][N]), no activation.  Category = first argmax of the logits starting
 * from -10000 (check_acc, general.hpp:112-119).  Nin == 0 / Nout == 0 drop the layer (the
 * category is then any(y > 0)). */
static void dense_forward(int rows, int fan_in, const float* w, const float* b, int clip_mode,
                          float clip, const float* in, float* out) {
  for (int o = 0; o < rows; ++o) {
    float acc = 0.f;
    for (int i = 0; i < fan_in; ++i) acc = fmaf(w[(size_t)o * fan_in + i], in[i], acc);
    const float z = acc + b[o];
    out[o] = clip_mode ? act(z, clip) : z;
  }
}

static int argmax_logits(const float* lg, int n) {
  float best = -10000.f;
  int arg = 0;
  for (int o = 0; o < n; ++o)
    if (lg[o] > best) {
      best = lg[o];
      arg = o;
    }
  return arg;
}

/* Plain model forward.  x0 is B x Nin (or B x N without an input layer); yL gets the last
 * hidden activations (B x N), logits B x Nout; all outputs nullable except cat. */
int oracle_model_plain(int Nin, const float* w_in, const float* b_in, float in_clip, int N, int L,
                       int K, const int32_t* cols, const float* vals, int bias_kind,
                       const float* bias, float clip, int Nout, const float* w_out,
                       const float* b_out, int64_t B, const float* x0, float* yL, float* logits,
                       int32_t* cat, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  sorted_net s;
  if (sort_net(N, L, K, cols, vals, &s)) return -1;
  const int width = Nin > 0 ? Nin : N;
  int rc = 0;
#pragma omp parallel
  {
    float* a = (float*)malloc((size_t)N * sizeof(float));
    float* b = (float*)malloc((size_t)N * sizeof(float));
    float* lg = (float*)malloc((size_t)(Nout > 0 ? Nout : 1) * sizeof(float));
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < B; ++i) {
      const float* x = x0 + (size_t)i * width;
      if (Nin > 0)
        dense_forward(N, Nin, w_in, b_in, 1, in_clip, x, a);
      else
        memcpy(a, x, (size_t)N * sizeof(float));
      for (int l = 0; l < L; ++l) {
        for (int o = 0; o < N; ++o)
          b[o] = act(row_dot(&s, l, o, a) + bias_of(bias_kind, bias, N, l, o), clip);
        float* t = a;
        a = b;
        b = t;
      }
      if (yL) memcpy(yL + (size_t)i * N, a, (size_t)N * sizeof(float));
      if (Nout > 0) {
        dense_forward(Nout, N, w_out, b_out, 0, 0.f, a, lg);
        if (logits) memcpy(logits + (size_t)i * Nout, lg, (size_t)Nout * sizeof(float));
        cat[i] = argmax_logits(lg, Nout);
      } else {
        int any = 0;
        for (int n = 0; n < N && !any; ++n) any = a[n] > 0.f;
        cat[i] = any;
      }
    }
    free(a);
    free(b);
    free(lg);
  }
  free_net(&s);
  return rc;
}

/* ---- lossy SNICIT (src/beyond/SNICIT/SNICIT.hpp:494-641, kernel.hpp:5-183) --------------
 * Every row is live.  At t, on the hidden activations Y_t:
 *   pool  = rows y*B/pool for y < pool (y_star_gen, kernel.hpp:12)
 *   prune in pool order: a present base i removes each other present row y with
 *         #{n : |P_y[n] - P_i[n]| > eps_pool} < N*eps_pool (kernel.hpp:19-41)
 *   m(i)  = first argmin_c #{n : |C_c[n] - Y_i[n]| > eps_assign}; centroids map to
 *           themselves (coarse_cluster, kernel.hpp:56-87)
 *   delta = |Y_i - C_m| > eps_assign ? Y_i - C_m : 0; non-empty iff any(delta != 0)
 *           (any(delta > 0) with positive_only, the reference's test, kernel.hpp:95)
 * Post layers (sparse_hidden_post + update_post, kernel.hpp:103-168):
 *   zc = W C + b, C' = f(zc);  y = f(zc[m] + W delta) - f(zc[m]), |y| <= eps_post -> 0,
 *   non-empty iff any(y != 0).
 * Recovery (kernel.hpp:170-183): Y_i = delta_i + C_m (C_m for centroids and empty rows), then
 * the output layer and argmax.  Outputs (nullable except cat): yL (recovered, B x N), logits,
 * pool_cols[pool], cent_cols[pool], assign[B], nerow_per_layer[L] (after each post layer),
 * info[4] = {B, pool, k, non-empty rows at t}. */
int oracle_model_snicit_lossy(int Nin, const float* w_in, const float* b_in, float in_clip, int N,
                              int L, int K, const int32_t* cols, const float* vals, int bias_kind,
                              const float* bias, float clip, int Nout, const float* w_out,
                              const float* b_out, int64_t B, const float* x0, int t, int pool,
                              float eps_pool, float eps_assign, float eps_post, int positive_only,
                              int32_t* cat, float* yL, float* logits, int32_t* pool_cols,
                              int32_t* cent_cols, int32_t* assign, int64_t* nerow_per_layer,
                              int64_t* info, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (t < 0 || t >= L || pool < 1) return -2;
  sorted_net s;
  if (sort_net(N, L, K, cols, vals, &s)) return -1;
  const int width = Nin > 0 ? Nin : N;
  float* Y = (float*)malloc((size_t)(B > 0 ? B : 1) * N * sizeof(float));
  /* input layer and pre-threshold layers, per sample */
#pragma omp parallel
  {
    float* b = (float*)malloc((size_t)N * sizeof(float));
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < B; ++i) {
      float* a = Y + (size_t)i * N;
      const float* x = x0 + (size_t)i * width;
      if (Nin > 0)
        dense_forward(N, Nin, w_in, b_in, 1, in_clip, x, a);
      else
        memcpy(a, x, (size_t)N * sizeof(float));
      for (int l = 0; l < t; ++l) {
        for (int o = 0; o < N; ++o)
          b[o] = act(row_dot(&s, l, o, a) + bias_of(bias_kind, bias, N, l, o), clip);
        memcpy(a, b, (size_t)N * sizeof(float));
      }
    }
    free(b);
  }
  const int P = B > 0 ? pool : 0;
  int32_t* pcol = (int32_t*)malloc((size_t)pool * sizeof(int32_t));
  int* alive = (int*)malloc((size_t)pool * sizeof(int));
  for (int y = 0; y < P; ++y) {
    pcol[y] = (int32_t)((int64_t)y * B / pool);
    alive[y] = 1;
  }
  const double eps_pool_d = decimal_of(eps_pool), eps_assign_d = decimal_of(eps_assign);
  const double eps_post_d = decimal_of(eps_post);
  const double limit = (double)N * eps_pool_d;
  for (int i = 0; i < P; ++i) {
    if (!alive[i]) continue;
    const float* pi = Y + (size_t)pcol[i] * N;
    for (int y = 0; y < P; ++y) {
      if (y == i || !alive[y]) continue;
      const float* py = Y + (size_t)pcol[y] * N;
      int cnt = 0;
      for (int n = 0; n < N; ++n) cnt += (double)fabsf(py[n] - pi[n]) > eps_pool_d;
      if ((double)cnt < limit) alive[y] = 0;
    }
  }
  int k = 0;
  int32_t* cent = (int32_t*)malloc((size_t)(pool > 0 ? pool : 1) * sizeof(int32_t));
  for (int y = 0; y < P; ++y)
    if (alive[y]) cent[k++] = pcol[y];
  float* C = (float*)malloc((size_t)(k > 0 ? k : 1) * N * sizeof(float));
  for (int c = 0; c < k; ++c) memcpy(C + (size_t)c * N, Y + (size_t)cent[c] * N, (size_t)N * sizeof(float));
  int32_t* m = (int32_t*)malloc((size_t)(B > 0 ? B : 1) * sizeof(int32_t));
  unsigned char* ne = (unsigned char*)calloc((size_t)(B > 0 ? B : 1), 1);
  int64_t nerow = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : nerow)
  for (int64_t i = 0; i < B; ++i) {
    float* y = Y + (size_t)i * N;
    int best = 0x7fffffff, bestc = 0;
    for (int c = 0; c < k; ++c) {
      const float* yc = C + (size_t)c * N;
      int d = 0;
      for (int n = 0; n < N; ++n) d += (double)fabsf(yc[n] - y[n]) > eps_assign_d;
      if (d < best) {
        best = d;
        bestc = c;
      }
    }
    int is_cent = 0;
    for (int c = 0; c < k && !is_cent; ++c)
      if (cent[c] == i) {
        is_cent = 1;
        bestc = c;
      }
    m[i] = bestc;
    if (is_cent || k == 0) continue;
    const float* yc = C + (size_t)bestc * N;
    int any = 0;
    for (int n = 0; n < N; ++n) {
      const float d = y[n] - yc[n];
      const float v = (double)fabsf(d) > eps_assign_d ? d : 0.f;
      y[n] = v;
      any |= positive_only ? (v > 0.f) : (v != 0.f);
    }
    ne[i] = (unsigned char)any;
    nerow += any;
  }
  if (info) {
    info[0] = B;
    info[1] = P;
    info[2] = k;
    info[3] = nerow;
  }
  if (pool_cols)
    for (int y = 0; y < P; ++y) pool_cols[y] = pcol[y];
  if (cent_cols)
    for (int c = 0; c < k; ++c) cent_cols[c] = cent[c];
  if (assign)
    for (int64_t i = 0; i < B; ++i) assign[i] = k > 0 ? m[i] : -1;
  float* Zc = (float*)malloc((size_t)(k > 0 ? k : 1) * N * sizeof(float));
  float* Cn = (float*)malloc((size_t)(k > 0 ? k : 1) * N * sizeof(float));
  for (int l = t; l < L; ++l) {
    for (int c = 0; c < k; ++c)
      for (int o = 0; o < N; ++o) {
        const float z = row_dot(&s, l, o, C + (size_t)c * N) + bias_of(bias_kind, bias, N, l, o);
        Zc[(size_t)c * N + o] = z;
        Cn[(size_t)c * N + o] = act(z, clip);
      }
    int64_t cnt = 0;
#pragma omp parallel reduction(+ : cnt)
    {
      float* outrow = (float*)malloc((size_t)N * sizeof(float));
#pragma omp for schedule(dynamic, 4)
      for (int64_t i = 0; i < B; ++i) {
        if (!ne[i]) continue;
        float* y = Y + (size_t)i * N;
        const float* zc = Zc + (size_t)m[i] * N;
        int any = 0;
        for (int o = 0; o < N; ++o) {
          const float z = row_dot(&s, l, o, y);
          float v = act(zc[o] + z, clip) - act(zc[o], clip);
          v = (double)fabsf(v) > eps_post_d ? v : 0.f;
          outrow[o] = v;
          any |= v != 0.f;
        }
        memcpy(y, outrow, (size_t)N * sizeof(float));
        ne[i] = (unsigned char)any;
        cnt += any;
      }
      free(outrow);
    }
    if (nerow_per_layer) nerow_per_layer[l] = cnt;
    float* sw = C;
    C = Cn;
    Cn = sw;
  }
  float* lg = (float*)malloc((size_t)(Nout > 0 ? Nout : 1) * sizeof(float));
  float* rec = (float*)malloc((size_t)N * sizeof(float));
  for (int64_t i = 0; i < B; ++i) {
    const float* yc = C + (size_t)(k > 0 ? m[i] : 0) * N;
    const float* y = Y + (size_t)i * N;
    for (int n = 0; n < N; ++n) rec[n] = k == 0 ? 0.f : (ne[i] ? y[n] + yc[n] : yc[n]);
    if (yL) memcpy(yL + (size_t)i * N, rec, (size_t)N * sizeof(float));
    if (Nout > 0) {
      dense_forward(Nout, N, w_out, b_out, 0, 0.f, rec, lg);
      if (logits) memcpy(logits + (size_t)i * Nout, lg, (size_t)Nout * sizeof(float));
      cat[i] = argmax_logits(lg, Nout);
    } else {
      int any = 0;
      for (int n = 0; n < N && !any; ++n) any = rec[n] > 0.f;
      cat[i] = any;
    }
  }
  free(lg);
  free(rec);
  free(Zc);
  free(Cn);
  free(C);
  free(m);
  free(ne);
  free(cent);
  free(alive);
  free(pcol);
  free(Y);
  free_net(&s);
  return 0;
}
