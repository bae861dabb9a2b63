/*
 * abi_harness.c -- a plain C client of include/ugx.h (built by
 * __graft_entry__.build() with gcc, linked against ugacc_amd/libugx.so).
 *
 *   abi_harness layout        sizeof / offsetof of every field of the ABI
 *                             structs, as one JSON line (no GPU needed)
 *   abi_harness solve DIR     CCSD(T) through the C ABI of the problem in
 *                             DIR (sizes.txt: "no nv"; fock.bin and the six
 *                             block files <name>.bin, little-endian FP64 in
 *                             the layouts of ugx.h), R_CONVERGENCE 1e-11,
 *                             then Lambda and the density energies; one
 *                             JSON line per step
 *
 * The tests compare the layout with the ctypes mirror in ugacc_amd/ugx.py
 * and the solve with a golden fixture.
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ugx.h"

#define FIELD(T, f) printf("%s\"%s\": [%zu, %zu]", first ? "" : ", ", #f, offsetof(T, f), sizeof(((T *)0)->f)), first = 0

static void layout(void) {
  int first = 1;
  printf("{\"ugx_options\": {\"size\": %zu, \"fields\": {", sizeof(ugx_options));
  FIELD(ugx_options, wfn_ccsd_t); FIELD(ugx_options, maxiter); FIELD(ugx_options, diis);
  FIELD(ugx_options, r_convergence); FIELD(ugx_options, ooc); FIELD(ugx_options, diis_nvec);
  FIELD(ugx_options, diis_min_vecs); FIELD(ugx_options, diis_policy); FIELD(ugx_options, precision);
  FIELD(ugx_options, polish_threshold); FIELD(ugx_options, guess); FIELD(ugx_options, energy_form);
  FIELD(ugx_options, print); FIELD(ugx_options, polish_keep_diis); FIELD(ugx_options, graph);
  FIELD(ugx_options, polish_mode);
  first = 1;
  printf("}}, \"ugx_result\": {\"size\": %zu, \"fields\": {", sizeof(ugx_result));
  FIELD(ugx_result, e_mp2); FIELD(ugx_result, e_ccsd); FIELD(ugx_result, e_ccsd_final); FIELD(ugx_result, e_t);
  FIELD(ugx_result, rms); FIELD(ugx_result, niter); FIELD(ugx_result, niter_low); FIELD(ugx_result, seconds_ccsd);
  FIELD(ugx_result, seconds_triples); FIELD(ugx_result, flops_ccsd); FIELD(ugx_result, device_allocs);
  first = 1;
  printf("}}, \"ugx_problem\": {\"size\": %zu, \"fields\": {", sizeof(ugx_problem));
  FIELD(ugx_problem, no); FIELD(ugx_problem, nv); FIELD(ugx_problem, fock); FIELD(ugx_problem, oooo);
  FIELD(ugx_problem, ooov); FIELD(ugx_problem, oovv); FIELD(ugx_problem, ovov); FIELD(ugx_problem, ovvv);
  FIELD(ugx_problem, vvvv);
  printf("}}}\n");
}

static double *read_block(const char *dir, const char *name, size_t n) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s.bin", dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  double *x = (double *)malloc(n * sizeof(double));
  size_t got = x ? fread(x, sizeof(double), n, f) : 0;
  fclose(f);
  if (got != n) {
    free(x);
    return NULL;
  }
  return x;
}

static int solve(const char *dir) {
  char path[4096];
  int no = 0, nv = 0;
  snprintf(path, sizeof(path), "%s/sizes.txt", dir);
  FILE *f = fopen(path, "r");
  if (!f || fscanf(f, "%d %d", &no, &nv) != 2 || no <= 0 || nv <= 0) {
    fprintf(stderr, "abi_harness: cannot read %s\n", path);
    return 2;
  }
  fclose(f);
  const size_t o = (size_t)no, v = (size_t)nv, n = o + v;
  const char *names[7] = {"fock", "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv"};
  const size_t sizes[7] = {n * n, o * o * o * o, o * o * o * v, o * o * v * v, o * v * o * v, o * v * v * v, v * v * v * v};
  double *b[7];
  for (int k = 0; k < 7; k++) {
    b[k] = read_block(dir, names[k], sizes[k]);
    if (!b[k]) {
      fprintf(stderr, "abi_harness: cannot read block %s\n", names[k]);
      return 2;
    }
  }
  ugx_ctx *ctx = NULL;
  int rc = ugx_create(&ctx, 0);
  if (rc != UGX_OK) {
    printf("{\"rc\": %d, \"stage\": \"create\"}\n", rc);
    return 1;
  }
  rc = ugx_load_hamiltonian(ctx, no, nv, b[0], b[1], b[2], b[3], b[4], b[5], b[6]);
  if (rc != UGX_OK) {
    printf("{\"rc\": %d, \"stage\": \"load\", \"error\": \"%s\"}\n", rc, ugx_last_error(ctx));
    ugx_destroy(ctx);
    return 1;
  }
  ugx_options opt;
  ugx_default_options(&opt);
  opt.wfn_ccsd_t = 1;
  opt.r_convergence = 1e-11;
  opt.maxiter = 200;
  ugx_result res;
  memset(&res, 0, sizeof(res));
  rc = ugx_ccsd(ctx, &opt, &res);
  printf("{\"rc\": %d, \"e_mp2\": %.17g, \"e_ccsd\": %.17g, \"e_t\": %.17g, \"niter\": %d, \"rms\": %.17g, "
         "\"device_allocs\": %d}\n",
         rc, res.e_mp2, res.e_ccsd, res.e_t, res.niter, res.rms, res.device_allocs);
  /* the post-CC steps of DERTYPE=FIRST for WFN=CCSD: Lambda, then the
   * density energies (ugacc.cc:118-137) */
  double pseudo = 0.0, lrms = 0.0, energies[8] = {0};
  int lniter = 0;
  int rcl = ugx_lambda(ctx, &opt, &pseudo, &lniter, &lrms, NULL);
  int rcd = rcl == UGX_OK ? ugx_density(ctx, energies, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL)
                          : rcl;
  printf("{\"rc_lambda\": %d, \"rc_density\": %d, \"pseudoenergy\": %.17g, \"niter_lambda\": %d, "
         "\"e_one\": %.17g, \"e_two\": %.17g}\n",
         rcl, rcd, pseudo, lniter, energies[0], energies[1]);
  /* a failure path through the ABI: MAXITER 1 is reported, not thrown */
  opt.maxiter = 1;
  int rc1 = ugx_ccsd(ctx, &opt, &res);
  printf("{\"rc_maxiter1\": %d, \"error\": \"%s\"}\n", rc1, ugx_last_error(ctx));
  ugx_destroy(ctx);
  for (int k = 0; k < 7; k++) free(b[k]);
  return rc == UGX_OK ? 0 : 1;
}

int main(int argc, char **argv) {
  if (argc >= 2 && strcmp(argv[1], "layout") == 0) {
    layout();
    return 0;
  }
  if (argc >= 3 && strcmp(argv[1], "solve") == 0) return solve(argv[2]);
  fprintf(stderr, "usage: abi_harness layout | solve DIR\n");
  return 2;
}
