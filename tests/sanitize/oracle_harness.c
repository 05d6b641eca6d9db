/* Runs every entry point of the CPU oracle (oracle/oracle.c) on small seeded nets, for
 * AddressSanitizer and UndefinedBehaviorSanitizer builds.  Exits 0 when they all return 0
 * and the exact SNICIT categories equal the plain ones. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int oracle_plain_f32(int N, int L, int K, const int32_t* cols, const float* vals, int bias_kind,
                     const float* bias, float clip, int64_t B, const float* y0, float* yL,
                     int32_t* cat, int64_t* live, int nthreads);
int oracle_plain_f64(int N, int L, int K, const int32_t* cols, const float* vals, int bias_kind,
                     const float* bias, float clip, int64_t B, const float* y0, double* yL,
                     int32_t* cat, int nthreads);
int oracle_layer_f64(int N, int K, const int32_t* cols_l, const float* vals_l, int bias_kind,
                     const float* bias_l, float clip, int64_t B, const float* yin, double* yout);
int oracle_snicit_exact_ex(int N, int L, int K, const int32_t* cols, const float* vals,
                           int bias_kind, const float* bias, float clip, int64_t B, const float* y0,
                           int t, int sample_size, int segments, float eps, int seg_layout,
                           int32_t* cat, float* yL, int32_t* pool_ids, float* F_out,
                           int32_t* cent_ids, int32_t* assign, int64_t* live_per_layer,
                           int64_t* nerow_per_layer, int64_t* info, int nthreads);
int oracle_model_plain(int Nin, const float* w_in, const float* b_in, float in_clip, int N, int L,
                       int K, const int32_t* cols, const float* vals, int bias_kind,
                       const float* bias, float clip, int Nout, const float* w_out,
                       const float* b_out, int64_t B, const float* x0, float* yL, float* logits,
                       int32_t* cat, int nthreads);
int oracle_model_snicit_lossy(int Nin, const float* w_in, const float* b_in, float in_clip, int N,
                              int L, int K, const int32_t* cols, const float* vals, int bias_kind,
                              const float* bias, float clip, int Nout, const float* w_out,
                              const float* b_out, int64_t B, const float* x0, int t, int pool,
                              float eps_pool, float eps_assign, float eps_post, int positive_only,
                              int32_t* cat, float* yL, float* logits, int32_t* pool_cols,
                              int32_t* cent_cols, int32_t* assign, int64_t* nerow_per_layer,
                              int64_t* info, int nthreads);

static uint64_t st = 88172645463325252ull;
static uint64_t rnd(void) {
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return st;
}
static float unit(void) { return (float)(rnd() % 1000000) / 1e6f; }

int main(void) {
  const int N = 1024, L = 6, K = 32, B = 70, Nin = 40, Nout = 10;
  int32_t* cols = malloc(sizeof(int32_t) * L * N * K);
  float* vals = malloc(sizeof(float) * L * N * K);
  for (int i = 0; i < L * N * K; ++i) {
    cols[i] = (int32_t)(rnd() % N);
    vals[i] = (i % 7 == 0) ? 0.f : 0.0625f;
  }
  float* y0 = calloc((size_t)B * N, sizeof(float));
  for (int b = 0; b < B; ++b)
    for (int n = 0; n < N; ++n) y0[b * N + n] = (b % 9 == 0) ? 0.f : (unit() < 0.2f ? 1.f : 0.f);
  memcpy(y0 + 40 * N, y0 + 3 * N, sizeof(float) * N);
  const float bias = -0.3f;
  float* yL = malloc(sizeof(float) * B * N);
  double* yd = malloc(sizeof(double) * B * N);
  int32_t cat[70], cat2[70], pool[64], cent[64], assign[70];
  int64_t live[6], nerow[6], info[4];
  float F[64 * 16];
  int rc = 0;
  rc |= oracle_plain_f32(N, L, K, cols, vals, 0, &bias, 32.f, B, y0, yL, cat, live, 2);
  rc |= oracle_plain_f64(N, L, K, cols, vals, 0, &bias, 32.f, B, y0, yd, cat2, 2);
  rc |= oracle_layer_f64(N, K, cols, vals, 0, &bias, 32.f, B, y0, yd);
  for (int layout = 0; layout < 2; ++layout)
    for (int t = 0; t < L; t += 2) {
      rc |= oracle_snicit_exact_ex(N, L, K, cols, vals, 0, &bias, 32.f, B, y0, t, 32, 16, 0.03f,
                                   layout, cat2, yL, pool, F, cent, assign, live, nerow, info, 2);
      for (int b = 0; b < B; ++b)
        if (cat2[b] != cat[b]) rc |= 1;
    }
  /* a model with dense ends and per-neuron biases, lossy clustering */
  const int M = 128, Km = 20;
  int32_t* mc = malloc(sizeof(int32_t) * L * M * Km);
  float* mv = malloc(sizeof(float) * L * M * Km);
  float* mb = malloc(sizeof(float) * L * M);
  for (int i = 0; i < L * M * Km; ++i) {
    mc[i] = (int32_t)(rnd() % M);
    mv[i] = unit() - 0.45f;
  }
  for (int i = 0; i < L * M; ++i) mb[i] = 0.05f * (unit() - 0.5f);
  float* w_in = malloc(sizeof(float) * M * Nin);
  float* b_in = malloc(sizeof(float) * M);
  float* w_out = malloc(sizeof(float) * Nout * M);
  float* b_out = malloc(sizeof(float) * Nout);
  for (int i = 0; i < M * Nin; ++i) w_in[i] = unit() - 0.5f;
  for (int i = 0; i < M; ++i) b_in[i] = 0.1f * unit();
  for (int i = 0; i < Nout * M; ++i) w_out[i] = unit() - 0.5f;
  for (int i = 0; i < Nout; ++i) b_out[i] = 0.f;
  float* x0 = malloc(sizeof(float) * B * Nin);
  for (int i = 0; i < B * Nin; ++i) x0[i] = unit();
  float* my = malloc(sizeof(float) * B * M);
  float* lg = malloc(sizeof(float) * B * Nout);
  int32_t pc[128], cc[128];
  rc |= oracle_model_plain(Nin, w_in, b_in, 1.f, M, L, Km, mc, mv, 2, mb, 1.f, Nout, w_out, b_out,
                           B, x0, my, lg, cat, 2);
  for (int t = 0; t < L; t += 3)
    rc |= oracle_model_snicit_lossy(Nin, w_in, b_in, 1.f, M, L, Km, mc, mv, 2, mb, 1.f, Nout, w_out,
                                    b_out, B, x0, t, 128, 0.03f, 0.04f, 0.05f, t & 1, cat2, my, lg,
                                    pc, cc, assign, nerow, info, 2);
  free(cols); free(vals); free(y0); free(yL); free(yd); free(mc); free(mv); free(mb);
  free(w_in); free(b_in); free(w_out); free(b_out); free(x0); free(my); free(lg);
  if (rc) {
    fprintf(stderr, "oracle harness: a call failed or SNICIT categories differ\n");
    return 1;
  }
  printf("oracle harness ok\n");
  return 0;
}
