/*
 * dqmc_harness.c -- drives libdqmc_hip.so through its C ABI alone (no Python), in the order
 * the reference's host code runs a simulation:
 *   DQMC(model; beta, ...)            dqmc_create (DQMCStack, stack.jl:113-206)
 *   init! / init_hopping_matrices     dqmc_set_hopping (DQMC.jl:144-150, stack.jl:209-249)
 *   field / RNG                       dqmc_set_field, dqmc_set_uniforms
 *   initialize_run                    dqmc_build_stack = reverse_build_stack + propagate
 *                                     (DQMC.jl:153-198)
 *   sweep_once! -> update(LocalSweep) dqmc_sweep(1) (DQMC.jl:200-225, local_updates.jl:82)
 *   calculate_greens(mc, slice)       dqmc_calculate_greens (stack.jl:525-583)
 * and checks every result against the committed config-A fixture (tests/golden/
 * config_a_raw.bin, written by tests/golden/make_golden.py from the CPU oracle): G after the
 * build and after the sweep within 1e-10 relative, the accept sequence bit for bit up to the
 * first proposal in the tie band |u - p| < 1e-12 (all of it, and the final field, when there
 * is none), G from scratch at slices 0, 13 and 40.
 *
 * Usage: dqmc_harness <config_a_raw.bin>; exit status 0 when every check holds.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/dqmc_hip.h"

static int failures = 0;

static void check(int ok, const char* what) {
  printf("%-58s %s\n", what, ok ? "ok" : "FAILED");
  if (!ok) failures++;
}

static double rel_maxdiff(const double* a, const double* b, size_t count) {
  double d = 0.0, m = 0.0;
  for (size_t i = 0; i < count; i++) {
    const double e = fabs(a[i] - b[i]);
    if (e > d || e != e) d = e;
    if (fabs(b[i]) > m) m = fabs(b[i]);
  }
  return d / (m > 0 ? m : 1.0);
}

static void* read_block(FILE* f, size_t bytes) {
  void* p = malloc(bytes ? bytes : 1);
  if (!p || fread(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "fixture truncated\n");
    exit(2);
  }
  return p;
}

#define RC(call)                                                                   \
  do {                                                                             \
    int rc_ = (call);                                                              \
    if (rc_ != DQMC_OK) {                                                          \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, dqmc_last_error(h));     \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s config_a_raw.bin\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) {
    perror(argv[1]);
    return 2;
  }
  int32_t hdr[3];
  double alpha;
  if (fread(hdr, sizeof hdr, 1, f) != 1 || fread(&alpha, sizeof alpha, 1, f) != 1) return 2;
  const int n = hdr[0], M = hdr[1], cnt = hdr[2];
  const size_t nn = (size_t)n * n, nm = (size_t)n * M;
  double* T = read_block(f, 8 * nn);
  double* G_start = read_block(f, 8 * nn);
  double* G_end = read_block(f, 8 * nn);
  double* G_s[3];
  for (int k = 0; k < 3; k++) G_s[k] = read_block(f, 8 * nn);
  int8_t* conf = read_block(f, nm);
  int8_t* final_conf = read_block(f, nm);
  double* u = read_block(f, 8 * (size_t)cnt);
  double* p_ref = read_block(f, 8 * (size_t)cnt);
  uint8_t* acc_ref = read_block(f, (size_t)cnt);
  fclose(f);
  printf("config A: N = %d, M = %d, alpha = %.17g\n", n, M, alpha);

  /* DQMC(AttractiveHubbardModel(L=4, U=4); beta = 4): one flavour, safe_mult 10 */
  dqmc_config_t cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.n_sites = n;
  cfg.flavors = 1;
  cfg.slices = M;
  cfg.safe_mult = 10;
  cfg.delta_tau = 0.1;
  cfg.alpha = alpha;
  cfg.precision = DQMC_PRECISION_FP64;
  cfg.check_propagation_error = 1;
  dqmc_chain_t h = NULL;
  RC(dqmc_create(&cfg, &h));
  RC(dqmc_set_hopping(h, T, n));
  RC(dqmc_set_field(h, conf, n));
  RC(dqmc_set_uniforms(h, u, cnt));
  RC(dqmc_set_decision_log(h, cnt));

  /* initialize_run: reverse_build_stack + propagate */
  RC(dqmc_build_stack(h));
  int cs, dir, cr;
  RC(dqmc_get_state(h, &cs, &dir, &cr));
  check(cs == 1 && dir == 1 && cr == 1, "state after the build (current_slice 1, up, range 1)");
  double* G = malloc(8 * nn);
  RC(dqmc_get_greens(h, G, n, 0));
  check(rel_maxdiff(G, G_start, nn) < 1e-10, "G after reverse_build_stack + propagate (1e-10)");

  /* sweep_once! -> update(LocalSweep): one sweep */
  int64_t accepted = 0;
  RC(dqmc_sweep(h, 1, &accepted));
  uint8_t* dec = malloc((size_t)cnt);
  double* p = malloc(8 * (size_t)cnt);
  int64_t len = 0;
  RC(dqmc_get_decision_log(h, dec, p, cnt, &len));
  check(len == cnt, "one decision per proposal (2 N M)");
  /* first proposal in the tie band, from the fixture's p and the uniforms in use order */
  int first_tie = cnt;
  for (int k = 0, j = 0; k < cnt; k++) {
    const double pk = p_ref[k];
    int tie;
    if (pk > 1.0) {
      tie = pk < 1.0 + 1e-12;
    } else {
      tie = fabs(u[j] - pk) < 1e-12;
      j++;
    }
    if (tie) { first_tie = k; break; }
  }
  int same = 1;
  for (int k = 0; k < first_tie && k < len; k++)
    if (dec[k] != acc_ref[k]) same = 0;
  check(same, "accept sequence bit-exact up to the first tie");
  int64_t acc_ref_sum = 0;
  for (int k = 0; k < cnt; k++) acc_ref_sum += acc_ref[k];
  if (first_tie == cnt) {
    int8_t* fc = malloc(nm);
    RC(dqmc_get_field(h, fc, n));
    check(memcmp(fc, final_conf, nm) == 0, "final field equal to the fixture's");
    check(accepted == acc_ref_sum, "accepted flips equal to the fixture's");
    RC(dqmc_get_greens(h, G, n, 0));
    check(rel_maxdiff(G, G_end, nn) < 1e-10, "G after the sweep (1e-10)");
    free(fc);
  } else {
    printf("tie at proposal %d: the rest of the sweep is not compared\n", first_tie);
  }
  RC(dqmc_get_state(h, &cs, &dir, &cr));
  check(cs == 1 && dir == 1, "state after a sweep (current_slice 1, up)");

  /* calculate_greens(mc, slice) from scratch on the fixture's field */
  RC(dqmc_set_field(h, conf, n));
  const int slices[3] = {0, 13, 40};
  for (int k = 0; k < 3; k++) {
    char what[80];
    RC(dqmc_calculate_greens(h, slices[k], G, n));
    snprintf(what, sizeof what, "calculate_greens(mc, %d) from scratch (1e-10)", slices[k]);
    check(rel_maxdiff(G, G_s[k], nn) < 1e-10, what);
  }
  dqmc_stats_t st;
  RC(dqmc_get_stats(h, &st));
  printf("stats: accepted %lld of %lld, uniforms %lld, ties %lld, stabilisations %lld, "
         "max propagation error %.3e, device errors %lld\n",
         (long long)st.accepted, (long long)st.proposed, (long long)st.uniforms_used,
         (long long)st.ties, (long long)st.stabilizations, st.propagation_error_max,
         (long long)st.device_errors);
  check(st.device_errors == 0, "no device errors");
  RC(dqmc_destroy(h));
  printf("%s\n", failures ? "FAILED" : "all checks passed");
  return failures ? 1 : 0;
}
