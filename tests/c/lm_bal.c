/* lm_bal.c -- a C host of the static library libfagra_gpu.a, through include/fagra_gpu.h only.
 *
 * Reads a BAL problem file (Problem::parse's grammar, examples/bal/model.rs:59-113), uploads it,
 * runs LevenbergMarquardt::<Schur>::solve_batch on the GPU (fg_lm_solve_batch) with the LM
 * defaults and prints the outcome as one line of key=value pairs.
 *
 * usage: lm_bal FILE PRECISION(0 = FP64, 1 = FP32) MAX_ITERATIONS [HUBER]
 */
#include <stdio.h>
#include <stdlib.h>

#include "fagra_gpu.h"

static char* read_file(const char* path, size_t* len) {
  FILE* f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char* buf = n >= 0 ? malloc((size_t)n + 1) : NULL;
  if (buf && fread(buf, 1, (size_t)n, f) != (size_t)n) {
    free(buf);
    buf = NULL;
  }
  fclose(f);
  if (buf) {
    buf[n] = 0;
    *len = (size_t)n;
  }
  return buf;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: lm_bal FILE PRECISION MAX_ITERATIONS [HUBER]\n");
    return 2;
  }
  int precision = atoi(argv[2]);
  size_t max_iterations = (size_t)strtoull(argv[3], NULL, 10);
  double huber = argc > 4 ? atof(argv[4]) : 0.0;
  size_t len = 0;
  char* text = read_file(argv[1], &len);
  if (!text) {
    fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  char err[256] = {0};
  fg_bal_sizes sz;
  fg_status st = fg_bal_scan(text, len, &sz, err, sizeof err);
  if (st != FG_OK) {
    fprintf(stderr, "parse: %s\n", err);
    return 3;
  }
  double* cams = malloc(sizeof(double) * 9 * sz.n_cameras);
  double* pts = malloc(sizeof(double) * 3 * sz.n_points);
  int64_t* oc = malloc(sizeof(int64_t) * sz.n_observations);
  int64_t* op = malloc(sizeof(int64_t) * sz.n_observations);
  double* px = malloc(sizeof(double) * 2 * sz.n_observations);
  st = fg_bal_parse(text, len, &sz, cams, pts, oc, op, px, err, sizeof err);
  free(text);
  if (st != FG_OK) {
    fprintf(stderr, "parse: %s\n", err);
    return 3;
  }
  fg_ba* ba = NULL;
  fg_schur* s = NULL;
  fg_schur_config cfg;
  fg_optimize_options opts;
  fg_lm_options lm;
  fg_lm_outcome out;
  st = fg_ba_create(sz.n_cameras, sz.n_points, sz.n_observations, cams, pts, oc, op, px, huber, precision, &ba);
  if (st == FG_OK) st = fg_schur_config_default(&cfg, precision);
  if (st == FG_OK) st = fg_schur_create(&cfg, &s);
  if (st == FG_OK) st = fg_optimize_options_default(&opts);
  if (st == FG_OK) st = fg_lm_options_default(&lm);
  if (st != FG_OK) {
    fprintf(stderr, "setup: %s\n", fg_status_string(st));
    return 4;
  }
  opts.max_iterations = max_iterations;
  opts.step_tolerance = 0.0;
  opts.cost_tolerance = 0.0;
  st = fg_lm_solve_batch(s, ba, &opts, &lm, NULL, NULL, &out);
  if (st == FG_OK || st == FG_NO_CONVERGENCE || st == FG_NO_PROGRESS) st = fg_ba_get_state(ba, cams, pts);
  printf("status=%d accepted=%zu rejected=%zu linearizations=%zu initial=%.17g cost=%.17g camera0=%.17g\n",
         out.status, out.accepted_steps, out.rejected_steps, out.linearizations, out.initial_cost, out.cost,
         cams[0]);
  fg_schur_destroy(s);
  fg_ba_destroy(ba);
  free(cams);
  free(pts);
  free(oc);
  free(op);
  free(px);
  return st == FG_DEVICE_ERROR ? 5 : 0;
}
