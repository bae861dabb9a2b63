/* Driver for the sanitized build of the oracle (tests/test_sanitizers.py):
 * every routine of oracle/ugacc_oracle.c on a small synthetic Hamiltonian
 * (o = 2, v = 4, positive-definite two-electron integrals from random
 * factors, f_ov != 0), built with -fsanitize=address,undefined, so that an
 * out-of-bounds access, a use after free, a leak or undefined behaviour in
 * any of them ends the run with an error.  Prints "OK" and the energies. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void *orc_create(int no, int nv, const double *fock, const double *ints);
void orc_destroy(void *p);
int orc_ccsd(void *p, int maxiter, double conv, int do_diis, int mode, int diis_policy, int min_vecs, double *out,
             double *trace);
double orc_energy_full(void *p);
double orc_tcorr(void *p);
double orc_tcorr_ooc(void *p);
double orc_tcorr_tjl(void *p);
void orc_build_tstar(void *p);
void orc_hbar(void *p, double *Hov, double *Hoo, double *Hvv, double *Hoooo, double *Hvvvv, double *Hvovv,
              double *Hooov, double *Hovvo, double *Hovov, double *Hvvvo, double *Hovoo);
void orc_lambda_init(void *p);
int orc_lambda(void *p, int maxiter, double conv, int do_diis, int diis_policy, int min_vecs, double *out,
               double *trace);
void orc_density(void *p, double *E, double *Doo, double *Dvv, double *Dov, double *Dvo, double *Goooo,
                 double *Gvvvv, double *Gooov, double *Gvvvo, double *Govov, double *Goovv);
void orc_tgrad(void *p, double *s1, double *s2, double *Doo, double *Dvv, double *Dov, double *Gooov, double *Gvvvo,
               double *Goovv);
void orc_tgrad_ooc(void *p, double *s1, double *s2, double *Doo, double *Dvv, double *Dov, double *Gooov,
                   double *Gvvvo, double *Goovv);
void orc_tgrad_enable(void *p, int on);
void *orc_pert_create(void *p, const double *pert, double omega);
void orc_pert_destroy(void *q);
int orc_pert_solve(void *q, int hand, int maxiter, double conv, int do_diis, int min_vecs, double *out);
double orc_linresp(void *qx, void *qy);

static unsigned long long rng_state = 88172645463325252ULL;
static double urand(void) { /* xorshift64, uniform in [-0.5, 0.5) */
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) / 9007199254740992.0 - 0.5;
}

static double *dalloc(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  if (!p) abort();
  return p;
}

int main(void) {
  const int no = 2, nv = 4, n = no + nv, rank = 8;
  const size_t n2 = (size_t)n * n, n4 = n2 * n2;
  /* (pq|rs) = sum_x B[pq][x] B[rs][x] with B symmetric in (p, q); <pq|rs> = (pr|qs) */
  double *B = dalloc(n2 * rank), *ints = dalloc(n4), *fock = dalloc(n2);
  for (int p = 0; p < n; p++)
    for (int q = 0; q <= p; q++)
      for (int x = 0; x < rank; x++) B[((size_t)p * n + q) * rank + x] = B[((size_t)q * n + p) * rank + x] = 0.3 * urand();
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++)
      for (int r = 0; r < n; r++)
        for (int s = 0; s < n; s++) {
          double v = 0.0;
          for (int x = 0; x < rank; x++) v += B[((size_t)p * n + r) * rank + x] * B[((size_t)q * n + s) * rank + x];
          ints[(((size_t)p * n + q) * n + r) * n + s] = v;
        }
  for (int i = 0; i < no; i++) fock[(size_t)i * n + i] = -1.0 + 0.2 * i;
  for (int a = 0; a < nv; a++) fock[(size_t)(no + a) * n + no + a] = 0.6 + 0.4 * a;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) fock[(size_t)i * n + no + a] = fock[(size_t)(no + a) * n + i] = 0.01 * urand();

  void *s = orc_create(no, nv, fock, ints);
  double out[8], trace[64];
  const int rc = orc_ccsd(s, 60, 1e-10, 1, 0, 0, 3, out, trace);
  orc_build_tstar(s);
  const double et = orc_tcorr(s), eo = orc_tcorr_ooc(s), ej = orc_tcorr_tjl(s), ecc = orc_energy_full(s);
  const size_t o = no, v = nv;
  double *H[11];
  const size_t hs[11] = {o * v, o * o, v * v, o * o * o * o, v * v * v * v, v * o * v * v,
                         o * o * o * v, o * v * v * o, o * v * o * v, v * v * v * o, o * v * o * o};
  for (int k = 0; k < 11; k++) H[k] = dalloc(hs[k]);
  orc_hbar(s, H[0], H[1], H[2], H[3], H[4], H[5], H[6], H[7], H[8], H[9], H[10]);
  const size_t tsz[8] = {o * v, o * o * v * v, o * o, v * v, o * v, o * o * o * v, v * v * v * o, o * o * v * v};
  double *T[8], *U[8];
  for (int k = 0; k < 8; k++) { T[k] = dalloc(tsz[k]); U[k] = dalloc(tsz[k]); }
  orc_tgrad_ooc(s, U[0], U[1], U[2], U[3], U[4], U[5], U[6], U[7]);
  orc_tgrad(s, T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7]);
  orc_tgrad_enable(s, 1);
  orc_lambda_init(s);
  double lout[8], ltrace[64];
  const int lrc = orc_lambda(s, 60, 1e-10, 1, 0, 3, lout, ltrace);
  double E[8];
  const size_t ds[10] = {o * o, v * v, o * v, v * o, o * o * o * o, v * v * v * v, o * o * o * v, v * v * v * o,
                         o * v * o * v, o * o * v * v};
  double *D[10];
  for (int k = 0; k < 10; k++) D[k] = dalloc(ds[k]);
  orc_density(s, E, D[0], D[1], D[2], D[3], D[4], D[5], D[6], D[7], D[8], D[9]);
  /* two perturbations: symmetric and antisymmetric one-electron operators */
  double *A = dalloc(n2), *C = dalloc(n2);
  for (int p = 0; p < n; p++)
    for (int q = 0; q <= p; q++) {
      A[(size_t)p * n + q] = A[(size_t)q * n + p] = urand();
      const double w = p == q ? 0.0 : urand();
      C[(size_t)p * n + q] = w;
      C[(size_t)q * n + p] = -w;
    }
  void *pa = orc_pert_create(s, A, 0.1), *pc = orc_pert_create(s, C, 0.1);
  double pout[4];
  int prc = 0;
  for (int hand = 0; hand < 2; hand++) {
    prc |= orc_pert_solve(pa, hand, 60, 1e-10, 1, 3, pout);
    prc |= orc_pert_solve(pc, hand, 60, 1e-10, 1, 3, pout);
  }
  const double lac = orc_linresp(pa, pc), lca = orc_linresp(pc, pa), laa = orc_linresp(pa, pa);
  orc_pert_destroy(pa);
  orc_pert_destroy(pc);
  orc_destroy(s);
  for (int k = 0; k < 11; k++) free(H[k]);
  for (int k = 0; k < 8; k++) { free(T[k]); free(U[k]); }
  for (int k = 0; k < 10; k++) free(D[k]);
  free(A); free(C); free(B); free(ints); free(fock);
  const double vals[] = {ecc, et, eo, ej, E[0], E[1], lac, lca, laa};
  for (size_t k = 0; k < sizeof(vals) / sizeof(vals[0]); k++)
    if (!isfinite(vals[k])) { printf("non-finite value %zu\n", k); return 1; }
  printf("OK ccsd rc %d E_CCSD %.12f (T) %.12f %.12f %.12f lambda rc %d E_one+E_two %.12f pert rc %d linresp %.10f %.10f %.10f\n",
         rc, ecc, et, eo, ej, lrc, E[0] + E[1], prc, lac, lca, laa);
  return 0;
}
