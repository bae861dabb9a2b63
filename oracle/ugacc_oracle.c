/*
 * ugacc_oracle.c -- FP64 CPU restatement of ugacc's closed-shell CCSD and (T)
 * hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only as the checker (or as the timed CPU baseline).  No product
 * code path calls it; the HIP library fails loudly instead of falling back.
 *
 * It keeps the reference's data layout (full nact^4 Dirac integral array,
 * <pq|rs> = ints[p][q][r][s], L = 2<pq|rs> - <pq|sr>), its loop structure and
 * its FP64 arithmetic, one function per reference routine, each citing the
 * reference file:line it follows (paths relative to the ugacc source tree).
 *
 * Pins (see tests/test_oracle_*.py): spin-orbital textbook CCSD/(T) on tiny
 * systems, brute-force FCI for two electrons, orbital-rotation invariance,
 * size extensivity, agreement of the three (T) variants, closed-form MP2,
 * and the known-answer energies in psi3/output.c1_fzc and
 * psi4/hof/output.{c1_nofzc,cs_fzc} through the chemistry front end.
 * The DIIS trajectory of Psi4's DIISManager is parity unpinned (third party,
 * absent); the attic/diis.cc Pulay DIIS restated here is pinned by the
 * iteration trace of psi3/output.c1_fzc:394-414.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int no, nv, nact;
  double *fock;            /* nact x nact */
  double *ints;            /* nact^4, <pq|rs> */
  double *L;               /* nact^4, 2<pq|rs> - <pq|sr> */
  double *D1, *D2;
  double *t1, *t1old, *t2, *t2old, *tau, *ttau;
  double *Fvv, *Foo, *Fov;
  double *Woooo, *Wovov, *Wovvo, *Wovoo;
  double *t1s, *t2s;
  /* DIIS storage (ring of up to 8 vectors of length no*nv + no^2 nv^2) */
  int diis_nvec;
  double *diis_err, *diis_amp, *diis_errnorm;
  /* upper bounds of the outermost occupied / virtual loops of build_F,
   * build_W and build_t1: no / nv except inside orc_iteration_sample */
  int lim_o, lim_v;
  int diis_count, diis_next;
  /* Lambda (cclambda.cc): HBAR of the amplitudes Lambda was started from,
   * l1/l2 and their previous values, Goo/Gvv, the half residual Z */
  double *hb[11];
  double *l1, *l1old, *l2, *l2old, *Goo, *Gvv, *lZ;
  /* (T) gradient terms of CCSD(T) with DERTYPE=FIRST (ccwfn.cc:854-1098):
   * the Lambda sources s1/s2 and the (T) density parts; tg_on: orc_lambda
   * and orc_density include them */
  int tg_on;
  double *tg_s1, *tg_s2, *tg_Doo, *tg_Dvv, *tg_Dov, *tg_Gooov, *tg_Gvvvo, *tg_Goovv;
} orc_state;

#define NA (s->nact)
#define I4(p, q, r, t) ((((size_t)(p) * NA + (q)) * NA + (r)) * NA + (t))
#define INT(p, q, r, t) (s->ints[I4(p, q, r, t)])
#define LL(p, q, r, t) (s->L[I4(p, q, r, t)])
#define F(p, q) (s->fock[(size_t)(p) * NA + (q)])
#define O2(i, a) ((size_t)(i) * nv + (a))
#define T4(i, j, a, b) ((((size_t)(i) * no + (j)) * nv + (a)) * nv + (b))
#define V3(a, b, c) ((((size_t)(a)) * nv + (b)) * nv + (c))

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* hamiltonian.cc:96-102 -- L(pqrs) = 2<pq|rs> - <pq|sr> */
static void build_L(orc_state *s) {
  int n = s->nact;
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++)
      for (int r = 0; r < n; r++)
        for (int t = 0; t < n; t++) LL(p, q, r, t) = 2 * INT(p, q, r, t) - INT(p, q, t, r);
}

/* ccwfn.cc:284-298 */
void orc_build_tau(orc_state *s) {
  int no = s->no, nv = s->nv;
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double x = s->t1[O2(i, a)] * s->t1[O2(j, b)];
          s->tau[T4(i, j, a, b)] = s->t2[T4(i, j, a, b)] + x;
          s->ttau[T4(i, j, a, b)] = s->t2[T4(i, j, a, b)] + 0.5 * x;
        }
}

/* ccwfn.cc:12-120 (denominators :62-72, guess :75-87, tau :91-94) */
void *orc_create(int no, int nv, const double *fock, const double *ints) {
  orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
  s->no = no; s->nv = nv; s->nact = no + nv;
  s->lim_o = no; s->lim_v = nv;
  size_t n = s->nact, n4 = n * n * n * n, ov = (size_t)no * nv, oovv = ov * ov;
  s->fock = zalloc(n * n); memcpy(s->fock, fock, n * n * sizeof(double));
  s->ints = zalloc(n4); memcpy(s->ints, ints, n4 * sizeof(double));
  s->L = zalloc(n4); build_L(s);
  s->D1 = zalloc(ov); s->D2 = zalloc(oovv);
  s->t1 = zalloc(ov); s->t1old = zalloc(ov);
  s->t2 = zalloc(oovv); s->t2old = zalloc(oovv);
  s->tau = zalloc(oovv); s->ttau = zalloc(oovv);
  s->Fvv = zalloc((size_t)nv * nv); s->Foo = zalloc((size_t)no * no); s->Fov = zalloc(ov);
  s->Woooo = zalloc((size_t)no * no * no * no);
  s->Wovov = zalloc(oovv); s->Wovvo = zalloc(oovv);
  s->Wovoo = zalloc(ov * no * no);
  s->t1s = zalloc(ov); s->t2s = zalloc(oovv);
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) s->D1[O2(i, a)] = F(i, i) - F(a + no, a + no);
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++)
          s->D2[T4(i, j, a, b)] = F(i, i) + F(j, j) - F(a + no, a + no) - F(b + no, b + no);
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) s->t1[O2(i, a)] = F(i, a + no) / s->D1[O2(i, a)];
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++)
          s->t2[T4(i, j, a, b)] = INT(i, j, a + no, b + no) / s->D2[T4(i, j, a, b)];
  orc_build_tau(s);
  return s;
}

static void diis_free(orc_state *s) {
  free(s->diis_err); free(s->diis_amp); free(s->diis_errnorm);
  s->diis_err = s->diis_amp = s->diis_errnorm = NULL;
  s->diis_count = s->diis_next = 0;
}

void orc_destroy(void *p) {
  orc_state *s = (orc_state *)p;
  if (!s) return;
  free(s->fock); free(s->ints); free(s->L); free(s->D1); free(s->D2);
  free(s->t1); free(s->t1old); free(s->t2); free(s->t2old); free(s->tau); free(s->ttau);
  free(s->Fvv); free(s->Foo); free(s->Fov);
  free(s->Woooo); free(s->Wovov); free(s->Wovvo); free(s->Wovoo);
  free(s->t1s); free(s->t2s);
  for (int k = 0; k < 11; k++) free(s->hb[k]);
  free(s->l1); free(s->l1old); free(s->l2); free(s->l2old); free(s->Goo); free(s->Gvv); free(s->lZ);
  free(s->tg_s1); free(s->tg_s2); free(s->tg_Doo); free(s->tg_Dvv); free(s->tg_Dov);
  free(s->tg_Gooov); free(s->tg_Gvvvo); free(s->tg_Goovv);
  diis_free(s);
  free(s);
}

void orc_set_amps(void *p, const double *t1, const double *t2) {
  orc_state *s = (orc_state *)p;
  size_t ov = (size_t)s->no * s->nv;
  memcpy(s->t1, t1, ov * sizeof(double));
  memcpy(s->t2, t2, ov * ov * sizeof(double));
  orc_build_tau(s);
}

void orc_get_amps(void *p, double *t1, double *t2) {
  orc_state *s = (orc_state *)p;
  size_t ov = (size_t)s->no * s->nv;
  if (t1) memcpy(t1, s->t1, ov * sizeof(double));
  if (t2) memcpy(t2, s->t2, ov * ov * sizeof(double));
}

void orc_get_tau(void *p, double *tau, double *ttau) {
  orc_state *s = (orc_state *)p;
  size_t ov = (size_t)s->no * s->nv;
  if (tau) memcpy(tau, s->tau, ov * ov * sizeof(double));
  if (ttau) memcpy(ttau, s->ttau, ov * ov * sizeof(double));
}

void orc_get_L(void *p, double *L) {
  orc_state *s = (orc_state *)p;
  size_t n = s->nact;
  memcpy(L, s->L, n * n * n * n * sizeof(double));
}

/* ccwfn.cc:260-282, reproducing the '=' at :272 (only the last (i,a) term of
 * the one-electron part survives).  fock_ov = 0 for a canonical RHF
 * reference, where this equals orc_energy_full. */
double orc_energy(void *p) {
  orc_state *s = (orc_state *)p;
  int no = s->no, nv = s->nv;
  double one = 0;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) one = 2 * F(i, a + no) * s->t1[O2(i, a)];
  double two = 0;
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) two += s->tau[T4(i, j, a, b)] * LL(i, j, a + no, b + no);
  return one + two;
}

/* ccwfn.cc:260-282 with the one-electron sum taken over all (i,a). */
double orc_energy_full(void *p) {
  orc_state *s = (orc_state *)p;
  int no = s->no, nv = s->nv;
  double one = 0;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) one += 2 * F(i, a + no) * s->t1[O2(i, a)];
  double two = 0;
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) two += s->tau[T4(i, j, a, b)] * LL(i, j, a + no, b + no);
  return one + two;
}

/* ccwfn.cc:300-309 */
static void amp_save(orc_state *s) {
  double *x = s->t2; s->t2 = s->t2old; s->t2old = x;
  double *y = s->t1; s->t1 = s->t1old; s->t1old = y;
}

/* ccwfn.cc:311-356 */
static void build_F(orc_state *s) {
  int no = s->no, nv = s->nv;
  const double *t1 = s->t1old, *ttau = s->ttau;
#pragma omp parallel for collapse(2) schedule(static)
  for (int a = 0; a < s->lim_v; a++)
    for (int e = 0; e < nv; e++) {
      double value = F(a + no, e + no);
      for (int m = 0; m < no; m++) {
        value -= 0.5 * F(m, e + no) * t1[O2(m, a)];
        for (int f = 0; f < nv; f++) {
          value += t1[O2(m, f)] * LL(m, a + no, f + no, e + no);
          for (int n = 0; n < no; n++) value -= ttau[T4(m, n, a, f)] * LL(m, n, e + no, f + no);
        }
      }
      s->Fvv[(size_t)a * nv + e] = value;
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < s->lim_o; m++)
    for (int i = 0; i < no; i++) {
      double value = F(m, i);
      for (int e = 0; e < nv; e++) {
        value += 0.5 * t1[O2(i, e)] * F(m, e + no);
        for (int n = 0; n < no; n++) {
          value += t1[O2(n, e)] * LL(m, n, i, e + no);
          for (int f = 0; f < nv; f++) value += ttau[T4(i, n, e, f)] * LL(m, n, e + no, f + no);
        }
      }
      s->Foo[(size_t)m * no + i] = value;
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < s->lim_o; m++)
    for (int e = 0; e < nv; e++) {
      double value = F(m, e + no);
      for (int n = 0; n < no; n++)
        for (int f = 0; f < nv; f++) value += t1[O2(n, f)] * LL(m, n, e + no, f + no);
      s->Fov[O2(m, e)] = value;
    }
}

/* ccwfn.cc:358-430 */
static void build_W(orc_state *s) {
  int no = s->no, nv = s->nv;
  const double *tau = s->tau, *t1 = s->t1old, *t2 = s->t2old;
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < s->lim_o; m++)
    for (int n = 0; n < no; n++)
      for (int i = 0; i < no; i++)
        for (int j = 0; j < no; j++) {
          double value = INT(m, n, i, j);
          for (int e = 0; e < nv; e++) {
            value += t1[O2(j, e)] * INT(m, n, i, e + no) + t1[O2(i, e)] * INT(m, n, e + no, j);
            for (int f = 0; f < nv; f++) value += tau[T4(i, j, e, f)] * INT(m, n, e + no, f + no);
          }
          s->Woooo[(((size_t)m * no + n) * no + i) * no + j] = value;
        }
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < s->lim_o; m++)
    for (int b = 0; b < nv; b++)
      for (int j = 0; j < no; j++)
        for (int e = 0; e < nv; e++) {
          double value = -INT(m, b + no, j, e + no);
          for (int f = 0; f < nv; f++) value -= t1[O2(j, f)] * INT(m, b + no, f + no, e + no);
          for (int n = 0; n < no; n++) value += t1[O2(n, b)] * INT(m, n, j, e + no);
          for (int n = 0; n < no; n++)
            for (int f = 0; f < nv; f++)
              value += INT(m, n, f + no, e + no) * (0.5 * t2[T4(j, n, f, b)] + t1[O2(j, f)] * t1[O2(n, b)]);
          s->Wovov[(((size_t)m * nv + b) * no + j) * nv + e] = value;
        }
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < s->lim_o; m++)
    for (int b = 0; b < nv; b++)
      for (int e = 0; e < nv; e++)
        for (int j = 0; j < no; j++) {
          double value = INT(m, b + no, e + no, j);
          for (int f = 0; f < nv; f++) value += t1[O2(j, f)] * INT(m, b + no, e + no, f + no);
          for (int n = 0; n < no; n++) value -= t1[O2(n, b)] * INT(m, n, e + no, j);
          for (int n = 0; n < no; n++)
            for (int f = 0; f < nv; f++)
              value -= INT(m, n, e + no, f + no) * (0.5 * t2[T4(j, n, f, b)] + t1[O2(j, f)] * t1[O2(n, b)]);
          for (int n = 0; n < no; n++)
            for (int f = 0; f < nv; f++) value += 0.5 * LL(m, n, e + no, f + no) * t2[T4(n, j, f, b)];
          s->Wovvo[(((size_t)m * nv + b) * nv + e) * no + j] = value;
        }
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < s->lim_o; m++)
    for (int b = 0; b < nv; b++)
      for (int i = 0; i < no; i++)
        for (int j = 0; j < no; j++) {
          double value = 0.0;
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) value += INT(m, b + no, e + no, f + no) * tau[T4(i, j, e, f)];
          s->Wovoo[(((size_t)m * nv + b) * no + i) * no + j] = value;
        }
}

/* ccwfn.cc:432-469 */
static void build_t1(orc_state *s) {
  int no = s->no, nv = s->nv;
  double *t1new = s->t1;
  const double *t1 = s->t1old, *t2 = s->t2old;
  const double *Fae = s->Fvv, *Fmi = s->Foo, *Fme = s->Fov;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < s->lim_o; i++)
    for (int a = 0; a < nv; a++) {
      double value = F(a + no, i);
      for (int e = 0; e < nv; e++) value += t1[O2(i, e)] * Fae[(size_t)a * nv + e];
      for (int m = 0; m < no; m++) value -= t1[O2(m, a)] * Fmi[(size_t)m * no + i];
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) value += (2 * t2[T4(i, m, a, e)] - t2[T4(i, m, e, a)]) * Fme[O2(m, e)];
      for (int n = 0; n < no; n++)
        for (int f = 0; f < nv; f++) value += t1[O2(n, f)] * LL(n, a + no, f + no, i);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++)
          for (int f = 0; f < nv; f++)
            value += (2 * t2[T4(m, i, e, f)] - t2[T4(m, i, f, e)]) * INT(m, a + no, e + no, f + no);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++)
          for (int n = 0; n < no; n++) value -= t2[T4(m, n, a, e)] * LL(n, m, e + no, i);
      t1new[O2(i, a)] = value;
    }
}

/* ccwfn.cc:471-539; rows i in [i0, i1) (the whole range for a normal
 * iteration; a sub-range is used only by the bounded CPU-baseline sample). */
static void build_t2_rows(orc_state *s, int i0, int i1) {
  int no = s->no, nv = s->nv;
  double *t2new = s->t2;
  const double *t2 = s->t2old, *tau = s->tau, *t1 = s->t1old;
  const double *Wmnij = s->Woooo, *Wmbej = s->Wovvo, *Wmbje = s->Wovov, *Zmbij = s->Wovoo;
  const double *Fme = s->Fov, *Fmi = s->Foo, *Fae = s->Fvv;
#define WOOOO(m, n, i, j) Wmnij[(((size_t)(m) * no + (n)) * no + (i)) * no + (j)]
#define WOVVO(m, b, e, j) Wmbej[(((size_t)(m) * nv + (b)) * nv + (e)) * no + (j)]
#define WOVOV(m, b, j, e) Wmbje[(((size_t)(m) * nv + (b)) * no + (j)) * nv + (e)]
#define ZOVOO(m, b, i, j) Zmbij[(((size_t)(m) * nv + (b)) * no + (i)) * no + (j)]
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = i0; i < i1; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double value = INT(i, j, a + no, b + no);
          for (int e = 0; e < nv; e++)
            value += t2[T4(i, j, a, e)] * Fae[(size_t)b * nv + e] + t2[T4(j, i, b, e)] * Fae[(size_t)a * nv + e];
          for (int e = 0; e < nv; e++)
            for (int m = 0; m < no; m++)
              value -= 0.5 * (t2[T4(i, j, a, e)] * t1[O2(m, b)] * Fme[O2(m, e)] +
                              t2[T4(i, j, e, b)] * t1[O2(m, a)] * Fme[O2(m, e)]);
          for (int m = 0; m < no; m++)
            value -= t2[T4(i, m, a, b)] * Fmi[(size_t)m * no + j] + t2[T4(m, j, a, b)] * Fmi[(size_t)m * no + i];
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              value -= 0.5 * (t2[T4(i, m, a, b)] * t1[O2(j, e)] * Fme[O2(m, e)] +
                              t2[T4(m, j, a, b)] * t1[O2(i, e)] * Fme[O2(m, e)]);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++) value += tau[T4(m, n, a, b)] * WOOOO(m, n, i, j);
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) value += tau[T4(i, j, e, f)] * INT(a + no, b + no, e + no, f + no);
          for (int e = 0; e < nv; e++)
            value += t1[O2(i, e)] * INT(a + no, b + no, e + no, j) + t1[O2(j, e)] * INT(b + no, a + no, e + no, i);
          for (int m = 0; m < no; m++)
            value -= t1[O2(m, a)] * INT(m, b + no, i, j) + t1[O2(m, b)] * INT(m, a + no, j, i);
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++) {
              value += (t2[T4(i, m, a, e)] - t2[T4(i, m, e, a)]) * WOVVO(m, b, e, j);
              value += (t2[T4(j, m, b, e)] - t2[T4(j, m, e, b)]) * WOVVO(m, a, e, i);
              value += t2[T4(i, m, a, e)] * (WOVVO(m, b, e, j) + WOVOV(m, b, j, e));
              value += t2[T4(j, m, b, e)] * (WOVVO(m, a, e, i) + WOVOV(m, a, i, e));
              value += t2[T4(m, j, a, e)] * WOVOV(m, b, i, e);
              value += t2[T4(i, m, e, b)] * WOVOV(m, a, j, e);
            }
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++) {
              value -= t1[O2(i, e)] * t1[O2(m, a)] * INT(m, b + no, e + no, j);
              value -= t1[O2(i, e)] * t1[O2(m, b)] * INT(m, a + no, j, e + no);
              value -= t1[O2(j, e)] * t1[O2(m, a)] * INT(m, b + no, i, e + no);
              value -= t1[O2(j, e)] * t1[O2(m, b)] * INT(m, a + no, e + no, i);
            }
          for (int m = 0; m < no; m++) value -= t1[O2(m, a)] * ZOVOO(m, b, i, j) + t1[O2(m, b)] * ZOVOO(m, a, j, i);
          t2new[T4(i, j, a, b)] = value;
        }
#undef WOOOO
#undef WOVVO
#undef WOVOV
#undef ZOVOO
}

/* ccwfn.cc:541-553 */
double orc_t1norm(void *p) {
  orc_state *s = (orc_state *)p;
  int no = s->no, nv = s->nv;
  double diag = 0;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) diag += s->t1[O2(i, a)] * s->t1[O2(i, a)];
  return sqrt(diag / (2 * no));
}

/* ccwfn.cc:555-582: rms = ||R||_2 (raw residual), t = t_old + R/D */
static double increment_amps(orc_state *s) {
  int no = s->no, nv = s->nv;
  double r1 = 0, r2 = 0;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      r1 += s->t1[O2(i, a)] * s->t1[O2(i, a)];
      s->t1[O2(i, a)] = s->t1old[O2(i, a)] + s->t1[O2(i, a)] / s->D1[O2(i, a)];
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++) {
          size_t x = T4(i, j, a, b);
          r2 += s->t2[x] * s->t2[x];
          s->t2[x] = s->t2old[x] + s->t2[x] / s->D2[x];
        }
    }
  return sqrt(r1 + r2);
}

/* One Jacobi step of ccwfn.cc:210-215 (amp_save, build_F, build_W,
 * build_t1, build_t2, increment_amps).  If R1/R2 are non-NULL the raw
 * residuals (before the denominator update) are copied out. */
double orc_iterate(void *p, double *R1, double *R2) {
  orc_state *s = (orc_state *)p;
  size_t ov = (size_t)s->no * s->nv;
  amp_save(s);
  build_F(s);
  build_W(s);
  build_t1(s);
  build_t2_rows(s, 0, s->no);
  if (R1) memcpy(R1, s->t1, ov * sizeof(double));
  if (R2) memcpy(R2, s->t2, ov * ov * sizeof(double));
  return increment_amps(s);
}

void orc_get_intermediates(void *p, double *Fvv, double *Foo, double *Fov, double *Woooo, double *Wovov,
                           double *Wovvo, double *Wovoo) {
  orc_state *s = (orc_state *)p;
  size_t no = s->no, nv = s->nv;
  if (Fvv) memcpy(Fvv, s->Fvv, nv * nv * sizeof(double));
  if (Foo) memcpy(Foo, s->Foo, no * no * sizeof(double));
  if (Fov) memcpy(Fov, s->Fov, no * nv * sizeof(double));
  if (Woooo) memcpy(Woooo, s->Woooo, no * no * no * no * sizeof(double));
  if (Wovov) memcpy(Wovov, s->Wovov, no * nv * no * nv * sizeof(double));
  if (Wovvo) memcpy(Wovvo, s->Wovvo, no * nv * nv * no * sizeof(double));
  if (Wovoo) memcpy(Wovoo, s->Wovoo, no * nv * no * no * sizeof(double));
}

/* ---------------------------------------------------------------------
 * Pulay DIIS, attic/diis.cc:13-161.  Error vector = t - t_old over
 * (t1 | t2) flattened in [i][a] and [i][j][a][b] order (attic/diis.cc:40-53);
 * B = error dot products with a -1 border, scaled by max |B_pq|
 * (:95-119); C solves B C = (0,..,0,-1) (:122-126); the new amplitudes are
 * the C-weighted sum of the stored amplitude vectors (:128-153).
 * policy 0: ring replacement (slot = count mod nvec), as attic/diis.cc:37-38.
 * policy 1: replace the vector with the largest error norm once full.
 * ------------------------------------------------------------------- */
static int lin_solve(double *A, double *b, int n) {
  /* Gaussian elimination with partial pivoting, A row-major n x n. */
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(A[r * n + c]) > fabs(A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0.0) return -1;
    if (piv != c) {
      for (int k = 0; k < n; k++) { double t = A[c * n + k]; A[c * n + k] = A[piv * n + k]; A[piv * n + k] = t; }
      double t = b[c]; b[c] = b[piv]; b[piv] = t;
    }
    for (int r = c + 1; r < n; r++) {
      double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; r--) {
    double x = b[r];
    for (int k = r + 1; k < n; k++) x -= A[r * n + k] * b[k];
    b[r] = x / A[r * n + r];
  }
  return 0;
}

/* Adds (t - t_old) and t to the subspace; if the subspace holds at least
 * min_vecs vectors, overwrites t1/t2 with the extrapolated amplitudes.
 * Returns the subspace size used (0 when no extrapolation happened). */
/* One Pulay step over the vector pair (x1 | x2) with their previous values
 * (x1o | x2o): the CCSD amplitudes (t) or the Lambda amplitudes (l), which
 * share the subspace storage (never both at once). */
static int diis_step_x(orc_state *s, double *x1, const double *x1o, double *x2, const double *x2o, int nvec,
                       int policy, int min_vecs) {
  size_t ov = (size_t)s->no * s->nv, len = ov + ov * ov;
  if (!s->diis_err) {
    s->diis_nvec = nvec;
    s->diis_err = zalloc(len * nvec);
    s->diis_amp = zalloc(len * nvec);
    s->diis_errnorm = zalloc(nvec);
    s->diis_count = 0; s->diis_next = 0;
  }
  int slot;
  if (s->diis_count < nvec) slot = s->diis_count;
  else if (policy == 0) slot = s->diis_next;
  else {
    slot = 0;
    for (int p = 1; p < nvec; p++)
      if (s->diis_errnorm[p] > s->diis_errnorm[slot]) slot = p;
  }
  double *e = s->diis_err + slot * len, *x = s->diis_amp + slot * len;
  double nrm = 0;
  for (size_t k = 0; k < ov; k++) { e[k] = x1[k] - x1o[k]; x[k] = x1[k]; nrm += e[k] * e[k]; }
  for (size_t k = 0; k < ov * ov; k++) {
    e[ov + k] = x2[k] - x2o[k]; x[ov + k] = x2[k]; nrm += e[ov + k] * e[ov + k];
  }
  s->diis_errnorm[slot] = nrm;
  if (s->diis_count < nvec) s->diis_count++;
  s->diis_next = (slot + 1) % nvec;
  int n = s->diis_count;
  if (n < min_vecs) return 0;
  double *B = zalloc((size_t)(n + 1) * (n + 1)), *C = zalloc(n + 1);
  for (int p = 0; p < n; p++)
    for (int q = 0; q <= p; q++) {
      double d = 0;
      const double *ep = s->diis_err + p * len, *eq = s->diis_err + q * len;
      for (size_t k = 0; k < len; k++) d += ep[k] * eq[k];
      B[p * (n + 1) + q] = B[q * (n + 1) + p] = d;
    }
  for (int p = 0; p < n; p++) { B[p * (n + 1) + n] = -1; B[n * (n + 1) + p] = -1; }
  B[n * (n + 1) + n] = 0;
  double mx = fabs(B[0]);
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++)
      if (fabs(B[p * (n + 1) + q]) > mx) mx = fabs(B[p * (n + 1) + q]);
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++) B[p * (n + 1) + q] /= mx;
  C[n] = -1;
  if (lin_solve(B, C, n + 1) != 0) { free(B); free(C); return 0; }
  for (size_t k = 0; k < ov; k++) {
    double v = 0;
    for (int p = 0; p < n; p++) v += C[p] * s->diis_amp[p * len + k];
    x1[k] = v;
  }
  for (size_t k = 0; k < ov * ov; k++) {
    double v = 0;
    for (int p = 0; p < n; p++) v += C[p] * s->diis_amp[p * len + ov + k];
    x2[k] = v;
  }
  free(B); free(C);
  return n;
}

static int diis_step(orc_state *s, int nvec, int policy, int min_vecs) {
  return diis_step_x(s, s->t1, s->t1old, s->t2, s->t2old, nvec, policy, min_vecs);
}

/* The CCSD driver.
 * mode 0: ccwfn.cc:188-231 -- energy evaluated after DIIS and build_tau, so
 *         the reported E_CCSD lags the final Jacobi update by one step; the
 *         loop breaks on ||R|| < conv before DIIS; the final t1/t2 are the
 *         un-extrapolated update; DIIS extrapolates once min_vecs vectors are
 *         stored (Psi4's "subspace_size() > 2" is min_vecs = 3).
 * mode 1: attic/ugacc.cc:100-114 -- energy of the fresh Jacobi update printed
 *         before the convergence test, DIIS after it (attic DIIS starts with
 *         two vectors, min_vecs = 2).
 * trace[0..maxiter] receives E_MP2 followed by the per-iteration energies.
 * Returns 0 on convergence, 1 if not converged (ccwfn.cc:230-231). */
int orc_ccsd(void *p, int maxiter, double conv, int do_diis, int mode, int diis_policy, int min_vecs,
             double *out /* emp2, eccsd, rms, niter */, double *trace) {
  orc_state *s = (orc_state *)p;
  diis_free(s);
  double emp2 = orc_energy(s), eccsd = emp2, rms = 0;
  if (trace) trace[0] = emp2;
  int iter, niter = 0;
  for (iter = 1; iter <= maxiter; iter++) {
    /* tau/ttau already hold the amplitudes being saved (the attic's
     * tau_build(iter, t1old, t2old) at attic/ugacc.cc:102 gives the same). */
    amp_save(s);
    build_F(s);
    build_W(s);
    build_t1(s);
    build_t2_rows(s, 0, s->no);
    rms = increment_amps(s);
    niter = iter;
    if (mode == 1) {
      orc_build_tau(s);
      eccsd = orc_energy(s);
      if (trace) trace[iter] = eccsd;
      if (rms < conv) break;
      if (do_diis) { diis_step(s, 8, diis_policy, min_vecs); }
      orc_build_tau(s);
    } else {
      if (rms < conv) break;
      if (do_diis) diis_step(s, 8, diis_policy, min_vecs);
      orc_build_tau(s);
      eccsd = orc_energy(s);
      if (trace) trace[iter] = eccsd;
    }
  }
  orc_build_tau(s);
  /* build_tstar, ccwfn.cc:628-645 */
  {
    int no = s->no, nv = s->nv;
    for (int i = 0; i < no; i++)
      for (int a = 0; a < nv; a++) {
        s->t1s[O2(i, a)] = 2.0 * s->t1[O2(i, a)];
        for (int j = 0; j < no; j++)
          for (int b = 0; b < nv; b++) s->t2s[T4(i, j, a, b)] = 4.0 * s->t2[T4(i, j, a, b)] - 2.0 * s->t2[T4(i, j, b, a)];
      }
  }
  if (out) { out[0] = emp2; out[1] = eccsd; out[2] = rms; out[3] = niter; }
  return rms < conv ? 0 : 1;
}

/* build_tstar, ccwfn.cc:628-645 (for callers that set amplitudes directly) */
void orc_build_tstar(void *p) {
  orc_state *s = (orc_state *)p;
  int no = s->no, nv = s->nv;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      s->t1s[O2(i, a)] = 2.0 * s->t1[O2(i, a)];
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++) s->t2s[T4(i, j, a, b)] = 4.0 * s->t2[T4(i, j, a, b)] - 2.0 * s->t2[T4(i, j, b, a)];
    }
}

/* ------------------------------- (T) ------------------------------------ */

/* The connected W3 sum of ccwfn.cc:1388-1418 (W3_ijk), also the numerator of
 * t3_ijk (:1318-1352) and of tcorr (:664-689). */
static inline double w3_elem(const orc_state *s, const double *t2, int i, int j, int k, int a, int b, int c) {
  int no = s->no, nv = s->nv;
  double value = 0.0;
  for (int e = 0; e < nv; e++) {
    value += +INT(i, e + no, a + no, b + no) * t2[T4(k, j, c, e)] + INT(i, e + no, a + no, c + no) * t2[T4(j, k, b, e)] +
             INT(k, e + no, c + no, a + no) * t2[T4(j, i, b, e)] + INT(k, e + no, c + no, b + no) * t2[T4(i, j, a, e)] +
             INT(j, e + no, b + no, c + no) * t2[T4(i, k, a, e)] + INT(j, e + no, b + no, a + no) * t2[T4(k, i, c, e)];
  }
  for (int m = 0; m < no; m++) {
    value -= +INT(j, k, m, c + no) * t2[T4(i, m, a, b)] + INT(k, j, m, b + no) * t2[T4(i, m, a, c)] +
             INT(i, j, m, b + no) * t2[T4(k, m, c, a)] + INT(j, i, m, a + no) * t2[T4(k, m, c, b)] +
             INT(k, i, m, a + no) * t2[T4(j, m, b, c)] + INT(i, k, m, c + no) * t2[T4(j, m, b, a)];
  }
  return value;
}

/* ccwfn.cc:647-743, in-core (T) with the full o^3 v^3 t3. */
double orc_tcorr(void *p) {
  orc_state *s = (orc_state *)p;
  int no = s->no, nv = s->nv;
  const double *t1 = s->t1, *t2 = s->t2;
  size_t vvv = (size_t)nv * nv * nv;
  double *t3 = zalloc((size_t)no * no * no * vvv);
#define T3(i, j, k, a, b, c) t3[((((size_t)(i) * no + (j)) * no + (k)) * vvv) + V3(a, b, c)]
#pragma omp parallel for collapse(3) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++)
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++) {
              double denom = F(i, i) + F(j, j) + F(k, k);
              denom -= F(a + no, a + no) + F(b + no, b + no) + F(c + no, c + no);
              T3(i, j, k, a, b, c) = w3_elem(s, t2, i, j, k, a, b, c) / denom;
            }
  double *X1 = zalloc((size_t)no * nv);
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++)
      for (int k = 0; k < no; k++)
        for (int l = 0; l < no; l++)
          for (int c = 0; c < nv; c++)
            for (int d = 0; d < nv; d++)
              X1[O2(i, a)] += (T3(i, k, l, a, c, d) - T3(l, k, i, a, c, d)) * LL(k, l, c + no, d + no);
  double *X2 = zalloc((size_t)no * no * nv * nv);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double x = 0;
          for (int k = 0; k < no; k++)
            for (int c = 0; c < nv; c++) x += (T3(i, j, k, a, b, c) - T3(k, j, i, a, b, c)) * F(k, c + no);
          for (int k = 0; k < no; k++)
            for (int c = 0; c < nv; c++)
              for (int l = 0; l < no; l++) {
                x -= T3(i, k, l, a, b, c) * LL(k, l, j, c + no);
                x += T3(l, k, i, a, b, c) * INT(k, l, j, c + no);
              }
          for (int k = 0; k < no; k++)
            for (int c = 0; c < nv; c++)
              for (int d = 0; d < nv; d++) {
                x += T3(i, j, k, a, c, d) * LL(b + no, k, c + no, d + no);
                x -= T3(k, j, i, a, c, d) * INT(b + no, k, c + no, d + no);
              }
          X2[T4(i, j, a, b)] = x;
        }
  double et = 0.0;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      et += t1[O2(i, a)] * X1[O2(i, a)];
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++) et += (2.0 * t2[T4(i, j, a, b)] - t2[T4(i, j, b, a)]) * X2[T4(i, j, a, b)];
    }
  et *= 2.0;
#undef T3
  free(t3); free(X1); free(X2);
  return et;
}

/* ccwfn.cc:745-793 with t3_ijk (:1318-1352): loop over all (i,j,k) with a
 * v^3 buffer; uses t1s/t2s from build_tstar. */
double orc_tcorr_ooc(void *p) {
  orc_state *s = (orc_state *)p;
  int no = s->no, nv = s->nv;
  const double *t2 = s->t2;
  size_t vvv = (size_t)nv * nv * nv;
  double *t3 = zalloc(vvv);
  double *X1 = zalloc((size_t)no * nv), *X2 = zalloc((size_t)no * no * nv * nv);
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++) {
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++) {
              double denom = F(i, i) + F(j, j) + F(k, k);
              denom -= F(a + no, a + no) + F(b + no, b + no) + F(c + no, c + no);
              t3[V3(a, b, c)] = w3_elem(s, t2, i, j, k, a, b, c) / denom;
            }
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++) {
              X1[O2(i, a)] += (t3[V3(a, b, c)] - t3[V3(c, b, a)]) * LL(j, k, b + no, c + no);
              X2[T4(i, j, a, b)] += (t3[V3(a, b, c)] - t3[V3(c, b, a)]) * F(k, c + no);
              double y = 2.0 * t3[V3(a, b, c)] - t3[V3(a, c, b)] - t3[V3(c, b, a)];
              for (int l = 0; l < no; l++) X2[T4(i, l, a, b)] -= y * INT(j, k, l, c + no);
              for (int d = 0; d < nv; d++) X2[T4(i, j, a, d)] += y * INT(d + no, k, b + no, c + no);
            }
      }
  double et = 0.0;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      et += s->t1s[O2(i, a)] * X1[O2(i, a)];
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++) et += s->t2s[T4(i, j, a, b)] * X2[T4(i, j, a, b)];
    }
  free(t3); free(X1); free(X2);
  return et;
}

/* Energy of one triplet (i >= j >= k) in ccwfn.cc:795-852, W3 from
 * W3_ijk (:1388-1418).  Buffers W3, V3 are nv^3 each.
 * fov != 0 adds t2_ij^ab f_kc + t2_ik^ac f_jb + t2_jk^bc f_ia to V: the
 * f_kc t3 part of tcorr's X2 (ccwfn.cc:706-709), which the TJL routine
 * omits, has exactly the form of the <ij|ab> t1 disconnected term with
 * <pq|xy> -> t2_pq^xy and t1 -> f_ov (both have the pair symmetry
 * X_pq^xy = X_qp^yx).  With it the sum over triplets is the value tcorr
 * and tcorr_ooc return (pinned in tests/test_oracle_pins.py). */
static double tjl_triplet_x(const orc_state *s, int i, int j, int k, double *W3, double *V3, int fov) {
  int no = s->no, nv = s->nv;
  const double *t1 = s->t1, *t2 = s->t2;
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int c = 0; c < nv; c++) W3[V3(a, b, c)] = w3_elem(s, t2, i, j, k, a, b, c);
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int c = 0; c < nv; c++) {
        double d = INT(i, j, a + no, b + no) * t1[O2(k, c)] + INT(i, k, a + no, c + no) * t1[O2(j, b)] +
                   INT(j, k, b + no, c + no) * t1[O2(i, a)];
        if (fov)
          d += t2[T4(i, j, a, b)] * F(k, c + no) + t2[T4(i, k, a, c)] * F(j, b + no) +
               t2[T4(j, k, b, c)] * F(i, a + no);
        V3[V3(a, b, c)] = (W3[V3(a, b, c)] + d) / (1.0 + (a == b) + (a == c) + (b == c));
      }
  double et = 0.0;
  for (int a = 0; a < nv; a++)
    for (int b = 0; b <= a; b++)
      for (int c = 0; c <= b; c++) {
        double X = W3[V3(a, b, c)] * V3[V3(a, b, c)] + W3[V3(a, c, b)] * V3[V3(a, c, b)] +
                   W3[V3(b, a, c)] * V3[V3(b, a, c)] + W3[V3(b, c, a)] * V3[V3(b, c, a)] +
                   W3[V3(c, a, b)] * V3[V3(c, a, b)] + W3[V3(c, b, a)] * V3[V3(c, b, a)];
        double Y = V3[V3(a, b, c)] + V3[V3(b, c, a)] + V3[V3(c, a, b)];
        double Z = V3[V3(a, c, b)] + V3[V3(b, a, c)] + V3[V3(c, b, a)];
        double denom = F(i, i) + F(j, j) + F(k, k);
        denom -= F(a + no, a + no) + F(b + no, b + no) + F(c + no, c + no);
        et += ((Y - 2.0 * Z) * (W3[V3(a, b, c)] + W3[V3(b, c, a)] + W3[V3(c, a, b)]) +
               (Z - 2.0 * Y) * (W3[V3(a, c, b)] + W3[V3(b, a, c)] + W3[V3(c, b, a)]) + 3.0 * X) *
              (2.0 - ((i == j) + (i == k) + (j == k))) / denom;
      }
  return et;
}

static double tjl_triplet(const orc_state *s, int i, int j, int k, double *W3, double *V3) {
  return tjl_triplet_x(s, i, j, k, W3, V3, 0);
}

/* Number of i >= j >= k triplets, o(o+1)(o+2)/6. */
long orc_ntriplets(int no) { return (long)no * (no + 1) * (no + 2) / 6; }

/* Triplet index -> (i,j,k) in the loop order of ccwfn.cc:810-812. */
static void triplet_of(long t, int *pi, int *pj, int *pk) {
  int i = 0;
  while ((long)(i + 1) * (i + 2) * (i + 3) / 6 <= t) i++;
  t -= (long)i * (i + 1) * (i + 2) / 6;
  int j = 0;
  while ((long)(j + 1) * (j + 2) / 2 <= t) j++;
  t -= (long)j * (j + 1) / 2;
  *pi = i; *pj = j; *pk = (int)t;
}

/* ccwfn.cc:795-852 over triplets [t0, t1) of the i>=j>=k order (the full
 * range gives tcorr_ooc_TJL; sub-ranges serve the bounded CPU sample);
 * fov != 0 folds tcorr's f_ov t3 terms in (tjl_triplet_x). */
static double tjl_range(void *p, long t0, long t1, int fov) {
  orc_state *s = (orc_state *)p;
  size_t vvv = (size_t)s->nv * s->nv * s->nv;
  double et = 0.0;
#pragma omp parallel reduction(+ : et)
  {
    double *W3 = zalloc(vvv), *V3b = zalloc(vvv);
#pragma omp for schedule(dynamic, 1)
    for (long t = t0; t < t1; t++) {
      int i, j, k;
      triplet_of(t, &i, &j, &k);
      et += tjl_triplet_x(s, i, j, k, W3, V3b, fov);
    }
    free(W3); free(V3b);
  }
  return et;
}

double orc_tcorr_tjl_range(void *p, long t0, long t1) { return tjl_range(p, t0, t1, 0); }

/* Triplets [t0, t1) of the folded bracket: the full range gives tcorr's
 * value (ccwfn.cc:647-743) for any f_ov. */
double orc_tjl_fov_range(void *p, long t0, long t1) { return tjl_range(p, t0, t1, 1); }

double orc_tcorr_tjl(void *p) {
  orc_state *s = (orc_state *)p;
  return orc_tcorr_tjl_range(p, 0, orc_ntriplets(s->no));
}

/* Per-triplet energies in i>=j>=k order (for parity of the GPU triplet
 * kernel on small systems). */
void orc_tjl_per_triplet(void *p, double *e_out) {
  orc_state *s = (orc_state *)p;
  size_t vvv = (size_t)s->nv * s->nv * s->nv;
  long nt = orc_ntriplets(s->no);
  double *W3 = zalloc(vvv), *V3b = zalloc(vvv);
  for (long t = 0; t < nt; t++) {
    int i, j, k;
    triplet_of(t, &i, &j, &k);
    e_out[t] = tjl_triplet(s, i, j, k, W3, V3b);
  }
  free(W3); free(V3b);
}

/* Per-triplet energies of the TJL bracket with the f_ov t3 terms folded in
 * (tjl_triplet_x with fov): their sum is tcorr (ccwfn.cc:647-743). */
void orc_tjl_fov_per_triplet(void *p, double *e_out) {
  orc_state *s = (orc_state *)p;
  size_t vvv = (size_t)s->nv * s->nv * s->nv;
  long nt = orc_ntriplets(s->no);
#pragma omp parallel
  {
    double *W3 = zalloc(vvv), *V3b = zalloc(vvv);
#pragma omp for schedule(dynamic, 1)
    for (long t = 0; t < nt; t++) {
      int i, j, k;
      triplet_of(t, &i, &j, &k);
      e_out[t] = tjl_triplet_x(s, i, j, k, W3, V3b, 1);
    }
    free(W3); free(V3b);
  }
}

/* Folded per-triplet energies (tjl_triplet_x with fov) of the n triplets
 * idx[0..n) of the i>=j>=k order, in parallel over the list (sampled
 * parity at sizes where the whole (T) is too slow). */
void orc_tjl_fov_list(void *p, long n, const long *idx, double *e_out) {
  orc_state *s = (orc_state *)p;
  size_t vvv = (size_t)s->nv * s->nv * s->nv;
#pragma omp parallel
  {
    double *W3 = zalloc(vvv), *V3b = zalloc(vvv);
#pragma omp for schedule(dynamic, 1)
    for (long q = 0; q < n; q++) {
      int i, j, k;
      triplet_of(idx[q], &i, &j, &k);
      e_out[q] = tjl_triplet_x(s, i, j, k, W3, V3b, 1);
    }
    free(W3); free(V3b);
  }
}

/* W3_ijk for one triplet (ccwfn.cc:1388-1418), for kernel parity. */
void orc_w3(void *p, int i, int j, int k, double *W3) {
  orc_state *s = (orc_state *)p;
  int nv = s->nv;
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int c = 0; c < nv; c++) W3[V3(a, b, c)] = w3_elem(s, s->t2, i, j, k, a, b, c);
}

/* Bounded CPU sample of one CCSD iteration: build_F, build_W, build_t1 and
 * build_t2 with their outermost occupied loops cut to [0, rows) (the Fvv
 * virtual loop to the same fraction), so the time is rows/no of a full
 * iteration's.  The amplitudes are left as they were. */
void orc_iteration_sample(void *p, int rows) {
  orc_state *s = (orc_state *)p;
  if (rows < 1) rows = 1;
  if (rows > s->no) rows = s->no;
  s->lim_o = rows;
  s->lim_v = (int)(((long)s->nv * rows + s->no - 1) / s->no);
  amp_save(s);
  build_F(s);
  build_W(s);
  build_t1(s);
  build_t2_rows(s, 0, rows);
  amp_save(s);
  s->lim_o = s->no;
  s->lim_v = s->nv;
}

/* Residuals of the current amplitudes with the t2 residual restricted to
 * rows i in [0, rows): R1 (no*nv, full) and R2 rows (rows*no*nv*nv).  The
 * amplitudes are left as they were.  For parity checks at sizes where a full
 * oracle iteration is too slow. */
void orc_residual_rows(void *p, int rows, double *R1, double *R2) {
  orc_state *s = (orc_state *)p;
  int no = s->no, nv = s->nv;
  if (rows < 1) rows = 1;
  if (rows > no) rows = no;
  amp_save(s);
  build_F(s);
  build_W(s);
  build_t1(s);
  build_t2_rows(s, 0, rows);
  memcpy(R1, s->t1, (size_t)no * nv * sizeof(double));
  memcpy(R2, s->t2, (size_t)rows * no * nv * nv * sizeof(double));
  amp_save(s);
  orc_build_tau(s);
}

/* Threads of the OpenMP build's parallel loops (1 times the serial
 * reference loop nest on the same state). */
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* HBAR::HBAR, hbar.cc:8-228: the similarity-transformed Hamiltonian's
 * one- and two-body blocks from the current amplitudes (tau rebuilt from
 * them first), in the reference's layouts:
 *   Hov[m][e] Hoo[m][i] Hvv[a][e] Hoooo[m][n][i][j] Hvvvv[a][b][e][f]
 *   Hvovv[a][m][e][f] Hooov[m][n][i][e] Hovvo[m][b][e][j] Hovov[m][b][j][e]
 *   Hvvvo[a][b][e][i] Hovoo[m][b][i][j]
 * Loop by loop the reference's expressions (hbar.cc:38-223). */
#define O4(a, b, c, d, n2, n3, n4) ((((size_t)(a) * (n2) + (b)) * (n3) + (c)) * (n4) + (d))
void orc_hbar(void *p, double *Hov, double *Hoo, double *Hvv, double *Hoooo, double *Hvvvv, double *Hvovv,
              double *Hooov, double *Hovvo, double *Hovov, double *Hvvvo, double *Hovoo) {
  orc_state *s = (orc_state *)p;
  const int no = s->no, nv = s->nv;
  const double *t1 = s->t1, *t2 = s->t2, *tau;
  orc_build_tau(s);
  tau = s->tau;
#define T1(i, a) t1[O2(i, a)]
#define T2(i, j, a, b) t2[T4(i, j, a, b)]
#define TAU(i, j, a, b) tau[T4(i, j, a, b)]
  /* hbar.cc:38-45 */
  for (int m = 0; m < no; m++)
    for (int e = 0; e < nv; e++) {
      double v = F(m, e + no);
      for (int n = 0; n < no; n++)
        for (int f = 0; f < nv; f++) v += T1(n, f) * LL(m, n, e + no, f + no);
      Hov[O2(m, e)] = v;
    }
  /* :47-59 */
  for (int m = 0; m < no; m++)
    for (int i = 0; i < no; i++) {
      double v = F(m, i);
      for (int e = 0; e < nv; e++) {
        v += T1(i, e) * F(m, e + no);
        for (int n = 0; n < no; n++) {
          v += T1(n, e) * LL(m, n, i, e + no);
          for (int f = 0; f < nv; f++) v += TAU(i, n, e, f) * LL(m, n, e + no, f + no);
        }
      }
      Hoo[(size_t)m * no + i] = v;
    }
  /* :61-73 */
  for (int a = 0; a < nv; a++)
    for (int e = 0; e < nv; e++) {
      double v = F(a + no, e + no);
      for (int m = 0; m < no; m++) {
        v -= F(m, e + no) * T1(m, a);
        for (int f = 0; f < nv; f++) {
          v += T1(m, f) * LL(a + no, m, e + no, f + no);
          for (int n = 0; n < no; n++) v -= TAU(m, n, f, a) * LL(m, n, f + no, e + no);
        }
      }
      Hvv[(size_t)a * nv + e] = v;
    }
  /* :75-86 */
  for (int m = 0; m < no; m++)
    for (int n = 0; n < no; n++)
      for (int i = 0; i < no; i++)
        for (int j = 0; j < no; j++) {
          double v = INT(m, n, i, j);
          for (int e = 0; e < nv; e++) {
            v += 2.0 * T1(j, e) * INT(m, n, i, e + no);
            for (int f = 0; f < nv; f++) v += TAU(i, j, e, f) * INT(m, n, e + no, f + no);
          }
          Hoooo[O4(m, n, i, j, no, no, no)] = v;
        }
  /* :88-99 */
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int e = 0; e < nv; e++)
        for (int f = 0; f < nv; f++) {
          double v = INT(a + no, b + no, e + no, f + no);
          for (int m = 0; m < no; m++) {
            v -= 2.0 * T1(m, b) * INT(a + no, m, e + no, f + no);
            for (int n = 0; n < no; n++) v += TAU(m, n, a, b) * INT(m, n, e + no, f + no);
          }
          Hvvvv[O4(a, b, e, f, nv, nv, nv)] = v;
        }
  /* :101-108 */
  for (int a = 0; a < nv; a++)
    for (int m = 0; m < no; m++)
      for (int e = 0; e < nv; e++)
        for (int f = 0; f < nv; f++) {
          double v = INT(a + no, m, e + no, f + no);
          for (int n = 0; n < no; n++) v -= T1(n, a) * INT(n, m, e + no, f + no);
          Hvovv[O4(a, m, e, f, no, nv, nv)] = v;
        }
  /* :110-117 */
  for (int m = 0; m < no; m++)
    for (int n = 0; n < no; n++)
      for (int i = 0; i < no; i++)
        for (int e = 0; e < nv; e++) {
          double v = INT(m, n, i, e + no);
          for (int f = 0; f < nv; f++) v += T1(i, f) * INT(n, m, e + no, f + no);
          Hooov[O4(m, n, i, e, no, no, nv)] = v;
        }
  /* :119-134 */
  for (int m = 0; m < no; m++)
    for (int b = 0; b < nv; b++)
      for (int e = 0; e < nv; e++)
        for (int j = 0; j < no; j++) {
          double v = INT(m, b + no, e + no, j);
          for (int f = 0; f < nv; f++) v += T1(j, f) * INT(m, b + no, e + no, f + no);
          for (int n = 0; n < no; n++) v -= T1(n, b) * INT(m, n, e + no, j);
          for (int n = 0; n < no; n++)
            for (int f = 0; f < nv; f++) v -= INT(m, n, e + no, f + no) * TAU(n, j, b, f);
          for (int n = 0; n < no; n++)
            for (int f = 0; f < nv; f++) v += T2(n, j, f, b) * LL(m, n, e + no, f + no);
          Hovvo[O4(m, b, e, j, nv, nv, no)] = v;
        }
  /* :136-149 */
  for (int m = 0; m < no; m++)
    for (int b = 0; b < nv; b++)
      for (int j = 0; j < no; j++)
        for (int e = 0; e < nv; e++) {
          double v = INT(m, b + no, j, e + no);
          for (int f = 0; f < nv; f++) v += T1(j, f) * INT(b + no, m, e + no, f + no);
          for (int n = 0; n < no; n++) v -= T1(n, b) * INT(m, n, j, e + no);
          for (int n = 0; n < no; n++)
            for (int f = 0; f < nv; f++) v -= INT(n, m, e + no, f + no) * TAU(j, n, f, b);
          Hovov[O4(m, b, j, e, nv, no, nv)] = v;
        }
  /* :151-187 */
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int e = 0; e < nv; e++)
        for (int i = 0; i < no; i++) {
          double v = INT(a + no, b + no, e + no, i);
          for (int f = 0; f < nv; f++) v += T1(i, f) * INT(a + no, b + no, e + no, f + no);
          for (int m = 0; m < no; m++) v -= T1(m, b) * INT(a + no, m, e + no, i) + T1(m, a) * INT(b + no, m, i, e + no);
          for (int m = 0; m < no; m++)
            for (int f = 0; f < nv; f++)
              v -= TAU(i, m, f, a) * INT(m, b + no, e + no, f + no) + TAU(i, m, f, b) * INT(a + no, m, e + no, f + no);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++) v += TAU(m, n, a, b) * INT(m, n, e + no, i);
          for (int m = 0; m < no; m++) v -= F(m, e + no) * T2(m, i, a, b);
          for (int m = 0; m < no; m++)
            for (int f = 0; f < nv; f++) v += T2(m, i, f, b) * LL(a + no, m, e + no, f + no);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int f = 0; f < nv; f++)
                v += INT(m, n, e + no, f + no) *
                     (T1(i, f) * T2(m, n, a, b) + T1(m, a) * T2(i, n, f, b) + T1(n, b) * T2(m, i, a, f));
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int f = 0; f < nv; f++)
                v -= (T1(m, f) * T2(n, i, a, b) + T1(n, a) * T2(m, i, f, b)) * LL(m, n, f + no, e + no);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int f = 0; f < nv; f++) v += T1(i, f) * T1(m, a) * T1(n, b) * INT(m, n, e + no, f + no);
          Hvvvo[O4(a, b, e, i, nv, nv, no)] = v;
        }
  /* :189-223 */
  for (int m = 0; m < no; m++)
    for (int b = 0; b < nv; b++)
      for (int i = 0; i < no; i++)
        for (int j = 0; j < no; j++) {
          double v = INT(m, b + no, i, j);
          for (int e = 0; e < nv; e++) v += T1(j, e) * INT(m, b + no, i, e + no) + T1(i, e) * INT(b + no, m, j, e + no);
          for (int n = 0; n < no; n++) v -= T1(n, b) * INT(m, n, i, j);
          for (int n = 0; n < no; n++)
            for (int e = 0; e < nv; e++)
              v -= TAU(i, n, e, b) * INT(n, m, j, e + no) + TAU(j, n, e, b) * INT(m, n, i, e + no);
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) v += INT(m, b + no, e + no, f + no) * TAU(i, j, e, f);
          for (int e = 0; e < nv; e++) v += F(m, e + no) * T2(i, j, e, b);
          for (int n = 0; n < no; n++)
            for (int e = 0; e < nv; e++) v += LL(m, n, i, e + no) * T2(n, j, e, b);
          for (int e = 0; e < nv; e++)
            for (int n = 0; n < no; n++)
              for (int f = 0; f < nv; f++)
                v -= INT(m, n, e + no, f + no) *
                     (T1(j, f) * T2(i, n, e, b) + T1(i, e) * T2(j, n, f, b) + T1(n, b) * T2(i, j, e, f));
          for (int e = 0; e < nv; e++)
            for (int n = 0; n < no; n++)
              for (int f = 0; f < nv; f++)
                v += (T1(i, e) * T2(n, j, f, b) + T1(n, f) * T2(i, j, e, b)) * LL(m, n, e + no, f + no);
          for (int e = 0; e < nv; e++)
            for (int n = 0; n < no; n++)
              for (int f = 0; f < nv; f++) v -= T1(j, f) * T1(i, e) * T1(n, b) * INT(m, n, e + no, f + no);
          Hovoo[O4(m, b, i, j, nv, no, no)] = v;
        }
#undef T1
#undef T2
#undef TAU
}
#undef O4

/* ---------------------------------------------------------------- Lambda
 * CCLambda (cclambda.cc:13-289): the Lagrange multipliers of the CCSD energy,
 * iterated over the HBAR of fixed amplitudes.  The HBAR blocks are built once
 * (orc_lambda_init, from the amplitudes current at that call: the converged
 * t1/t2 after orc_ccsd) in the layouts of orc_hbar.  The CCSD_T/DERTYPE=FIRST
 * terms s1 and s2/2 (cclambda.cc:181, :243) come from orc_tgrad (ccwfn.cc:854-1098)
 * and are added when orc_tgrad_enable(1) is in effect. */
enum { HOV, HOO, HVV, HOOOO, HVVVV, HVOVV, HOOOV, HOVVO, HOVOV, HVVVO, HOVOO };
#define Q4(a, b, c, d, n2, n3, n4) ((((size_t)(a) * (n2) + (b)) * (n3) + (c)) * (n4) + (d))

/* cclambda.cc:28-42: l1 = 2 t1, l2 = 2 (2 t2 - t2^T); HBAR of the current
 * amplitudes into s->hb. */
void orc_lambda_init(void *p) {
  orc_state *s = (orc_state *)p;
  const size_t o = s->no, v = s->nv, ov = o * v, oovv = ov * ov;
  const size_t sz[11] = {o * v, o * o, v * v, o * o * o * o, v * v * v * v, v * o * v * v, o * o * o * v,
                         o * v * v * o, o * v * o * v, v * v * v * o, o * v * o * o};
  for (int k = 0; k < 11; k++)
    if (!s->hb[k]) s->hb[k] = zalloc(sz[k]);
  if (!s->l1) {
    s->l1 = zalloc(ov); s->l1old = zalloc(ov);
    s->l2 = zalloc(oovv); s->l2old = zalloc(oovv); s->lZ = zalloc(oovv);
    s->Goo = zalloc(o * o); s->Gvv = zalloc(v * v);
  }
  orc_hbar(s, s->hb[HOV], s->hb[HOO], s->hb[HVV], s->hb[HOOOO], s->hb[HVVVV], s->hb[HVOVV], s->hb[HOOOV],
           s->hb[HOVVO], s->hb[HOVOV], s->hb[HVVVO], s->hb[HOVOO]);
  const int no = s->no, nv = s->nv;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      s->l1[O2(i, a)] = 2.0 * s->t1[O2(i, a)];
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++)
          s->l2[T4(i, j, a, b)] = 2.0 * (2.0 * s->t2[T4(i, j, a, b)] - s->t2[T4(i, j, b, a)]);
    }
}

/* build_G, cclambda.cc:137-156, from l2 (the saved l2old in the reference):
 *   Goo[m][i] = sum_jab t2_mj^ab l2_ij^ab,  Gvv[a][e] = -sum_ijb t2_ij^eb l2_ij^ab */
static void lambda_G(orc_state *s, const double *l2) {
  const int no = s->no, nv = s->nv;
  const double *t2 = s->t2;
  for (int m = 0; m < no; m++)
    for (int i = 0; i < no; i++) {
      double value = 0.0;
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++)
          for (int j = 0; j < no; j++) value += t2[T4(m, j, a, b)] * l2[T4(i, j, a, b)];
      s->Goo[(size_t)m * no + i] = value;
    }
  for (int a = 0; a < nv; a++)
    for (int e = 0; e < nv; e++) {
      double value = 0.0;
      for (int i = 0; i < no; i++)
        for (int j = 0; j < no; j++)
          for (int b = 0; b < nv; b++) value -= t2[T4(i, j, e, b)] * l2[T4(i, j, a, b)];
      s->Gvv[(size_t)a * nv + e] = value;
    }
}

#define HB(k) (s->hb[k])
#define Hov_(m, e) HB(HOV)[O2(m, e)]
#define Hoo_(m, i) HB(HOO)[(size_t)(m) * no + (i)]
#define Hvv_(a, e) HB(HVV)[(size_t)(a) * nv + (e)]
#define Hoooo_(m, n, i, j) HB(HOOOO)[Q4(m, n, i, j, no, no, no)]
#define Hvvvv_(a, b, e, f) HB(HVVVV)[Q4(a, b, e, f, nv, nv, nv)]
#define Hvovv_(a, m, e, f) HB(HVOVV)[Q4(a, m, e, f, no, nv, nv)]
#define Hooov_(m, n, i, e) HB(HOOOV)[Q4(m, n, i, e, no, no, nv)]
#define Hovvo_(m, b, e, j) HB(HOVVO)[Q4(m, b, e, j, nv, nv, no)]
#define Hovov_(m, b, j, e) HB(HOVOV)[Q4(m, b, j, e, nv, no, nv)]
#define Hvvvo_(a, b, e, i) HB(HVVVO)[Q4(a, b, e, i, nv, nv, no)]
#define Hovoo_(m, b, i, j) HB(HOVOO)[Q4(m, b, i, j, nv, no, no)]
#define Goo_(m, i) s->Goo[(size_t)(m) * no + (i)]
#define Gvv_(a, e) s->Gvv[(size_t)(a) * nv + (e)]

/* build_l1, cclambda.cc:158-210: the raw l1 residual R1 from (l1, l2) and G */
static void lambda_l1(orc_state *s, const double *l1, const double *l2, double *R1) {
  const int no = s->no, nv = s->nv;
#define L1(i, a) l1[O2(i, a)]
#define L2(i, j, a, b) l2[T4(i, j, a, b)]
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      double value = 2 * Hov_(i, a);
      if (s->tg_on) value += s->tg_s1[O2(i, a)];  /* cclambda.cc:181 */
      for (int e = 0; e < nv; e++) value += L1(i, e) * Hvv_(e, a);
      for (int m = 0; m < no; m++) value -= L1(m, a) * Hoo_(i, m);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) value += L1(m, e) * (2 * Hovvo_(i, e, a, m) - Hovov_(i, e, m, a));
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++)
          for (int f = 0; f < nv; f++) value += L2(i, m, e, f) * Hvvvo_(e, f, a, m);
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++) value -= L2(m, n, a, e) * Hovoo_(i, e, m, n);
      for (int e = 0; e < nv; e++)
        for (int f = 0; f < nv; f++) value -= Gvv_(e, f) * (2 * Hvovv_(e, i, f, a) - Hvovv_(e, i, a, f));
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++) value -= Goo_(m, n) * (2 * Hooov_(m, i, n, a) - Hooov_(i, m, n, a));
      R1[O2(i, a)] = value;
    }
}

/* build_l2, cclambda.cc:212-268: Z from (l1, l2) and G, then the raw l2
 * residual R2_ijab = Z_ijab + Z_jiba */
static void lambda_l2(orc_state *s, const double *l1, const double *l2, double *R2) {
  const int no = s->no, nv = s->nv;
  double *Z = s->lZ;
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double value = LL(i, j, a + no, b + no);
          if (s->tg_on) value += 0.5 * s->tg_s2[T4(i, j, a, b)];  /* cclambda.cc:243 */
          value += 2.0 * L1(i, a) * Hov_(j, b) - L1(j, a) * Hov_(i, b);
          for (int e = 0; e < nv; e++) value += L2(i, j, e, b) * Hvv_(e, a);
          for (int m = 0; m < no; m++) value -= L2(m, j, a, b) * Hoo_(i, m);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++) value += 0.5 * L2(m, n, a, b) * Hoooo_(i, j, m, n);
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) value += 0.5 * L2(i, j, e, f) * Hvvvv_(e, f, a, b);
          for (int e = 0; e < nv; e++) value += L1(i, e) * (2 * Hvovv_(e, j, a, b) - Hvovv_(e, j, b, a));
          for (int m = 0; m < no; m++) value -= L1(m, b) * (2 * Hooov_(j, i, m, a) - Hooov_(i, j, m, a));
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++) {
              value += (2 * Hovvo_(i, e, a, m) - Hovov_(i, e, m, a)) * L2(m, j, e, b);
              value -= Hovov_(j, e, m, a) * L2(m, i, b, e);
              value -= Hovvo_(j, e, a, m) * L2(m, i, e, b);
            }
          for (int e = 0; e < nv; e++) value += Gvv_(a, e) * LL(i, j, e + no, b + no);
          for (int m = 0; m < no; m++) value -= Goo_(m, i) * LL(m, j, a + no, b + no);
          Z[T4(i, j, a, b)] = value;
        }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) R2[T4(i, j, a, b)] = Z[T4(i, j, a, b)] + Z[T4(j, i, b, a)];
}
#undef L1
#undef L2

/* pseudoenergy, cclambda.cc:270-288: 1/2 sum <ij|ab> l2_ij^ab */
static double lambda_pseudo(const orc_state *s, const double *l2) {
  const int no = s->no, nv = s->nv;
  double energy = 0.0;
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) energy += 0.5 * INT(i, j, a + no, b + no) * l2[T4(i, j, a, b)];
  return energy;
}

double orc_lambda_pseudoenergy(void *p) {
  orc_state *s = (orc_state *)p;
  return s->l2 ? lambda_pseudo(s, s->l2) : 0.0;
}

/* One evaluation of build_G, build_l1, build_l2 at (l1, l2) over the HBAR
 * built by orc_lambda_init: R1, R2 (raw, before the denominators) and the
 * G intermediates.  The context's own l1/l2 are not touched. */
void orc_lambda_residual(void *p, const double *l1, const double *l2, double *R1, double *R2, double *Goo,
                         double *Gvv) {
  orc_state *s = (orc_state *)p;
  lambda_G(s, l2);
  lambda_l1(s, l1, l2, R1);
  lambda_l2(s, l1, l2, R2);
  if (Goo) memcpy(Goo, s->Goo, (size_t)s->no * s->no * sizeof(double));
  if (Gvv) memcpy(Gvv, s->Gvv, (size_t)s->nv * s->nv * sizeof(double));
}

/* CCLambda::compute_lambda, cclambda.cc:93-135: from the guess of
 * orc_lambda_init, amp_save (:59-68), build_G/l1/l2, increment_amps
 * (:70-91: l = l_old + R/D with the CCSD denominators, rms the 2-norm of the
 * raw R), break on rms < conv before DIIS, else DIIS (the Pulay step of
 * orc_ccsd on (l1 | l2), extrapolating from min_vecs stored vectors) and the
 * pseudoenergy of the extrapolated l2.  trace[0..niter]: the pseudoenergy of
 * the guess, after each iteration's DIIS, and of the final l2 last.
 * out: pseudoenergy of the final l2, rms, niter.  Returns 0 on convergence,
 * 1 if not converged (cclambda.cc:133-134). */
int orc_lambda(void *p, int maxiter, double conv, int do_diis, int diis_policy, int min_vecs, double *out,
               double *trace) {
  orc_state *s = (orc_state *)p;
  const int no = s->no, nv = s->nv;
  const size_t ov = (size_t)no * nv;
  diis_free(s);
  if (trace) trace[0] = lambda_pseudo(s, s->l2);
  double rms = 0.0;
  int niter = 0;
  for (int iter = 1; iter <= maxiter; iter++) {
    double *x = s->l1; s->l1 = s->l1old; s->l1old = x;
    x = s->l2; s->l2 = s->l2old; s->l2old = x;
    orc_lambda_residual(s, s->l1old, s->l2old, s->l1, s->l2, NULL, NULL);
    double r1 = 0.0, r2 = 0.0;
    for (size_t k = 0; k < ov; k++) {
      r1 += s->l1[k] * s->l1[k];
      s->l1[k] = s->l1old[k] + s->l1[k] / s->D1[k];
    }
    for (size_t k = 0; k < ov * ov; k++) {
      r2 += s->l2[k] * s->l2[k];
      s->l2[k] = s->l2old[k] + s->l2[k] / s->D2[k];
    }
    rms = sqrt(r1 + r2);
    niter = iter;
    if (rms < conv) break;
    if (do_diis) diis_step_x(s, s->l1, s->l1old, s->l2, s->l2old, 8, diis_policy, min_vecs);
    if (trace) trace[iter] = lambda_pseudo(s, s->l2);
  }
  const double pe = lambda_pseudo(s, s->l2);
  if (trace) trace[niter] = pe;
  if (out) { out[0] = pe; out[1] = rms; out[2] = niter; }
  diis_free(s);
  return rms < conv ? 0 : 1;
}

void orc_get_lambda(void *p, double *l1, double *l2) {
  orc_state *s = (orc_state *)p;
  size_t ov = (size_t)s->no * s->nv;
  if (l1 && s->l1) memcpy(l1, s->l1, ov * sizeof(double));
  if (l2 && s->l2) memcpy(l2, s->l2, ov * ov * sizeof(double));
}

void orc_set_lambda(void *p, const double *l1, const double *l2) {
  orc_state *s = (orc_state *)p;
  size_t ov = (size_t)s->no * s->nv;
  if (!s->l1) orc_lambda_init(s);
  memcpy(s->l1, l1, ov * sizeof(double));
  memcpy(s->l2, l2, ov * ov * sizeof(double));
}
#undef HB

/* ------------------------------------------------------------- densities
 * CCDensity::onepdm / twopdm (ccdensity.cc:89-339): the CCSD one- and
 * two-particle density blocks from the amplitudes (t1, t2, tau as current)
 * and the Lambda amplitudes (l1, l2 of orc_lambda), and the energies they
 * give with the Fock matrix and the integrals.  With orc_tgrad_enable(1) the
 * blocks start from the (T) parts of orc_tgrad, as CCSD_T copies them in
 * first (ccdensity.cc:43-74); otherwise these are the CCSD densities.
 *   E[0] = E_one (ccdensity.cc:138-146), E[1] = E_two, E[2..7] = the
 *   OOOO, VVVV, OOOV, VVVO, OVOV, OOVV parts of E_two (the values printed).
 * Blocks (each may be NULL): Doo[i][j] Dvv[a][b] Dov[i][a] Dvo[a][i]
 * Goooo[i][j][k][l] Gvvvv[a][b][c][d] Gooov[i][j][k][a] Gvvvo[a][b][c][i]
 * Govov[i][a][j][b] Goovv[i][j][a][b]. */
void orc_density(void *p, double *E, double *Doo_o, double *Dvv_o, double *Dov_o, double *Dvo_o, double *Goooo_o,
                 double *Gvvvv_o, double *Gooov_o, double *Gvvvo_o, double *Govov_o, double *Goovv_o) {
  orc_state *s = (orc_state *)p;
  const int no = s->no, nv = s->nv;
  const size_t o = no, v = nv;
  orc_build_tau(s);
  const double *t1 = s->t1, *t2 = s->t2, *tau = s->tau, *l1 = s->l1, *l2 = s->l2;
  double *Doo = zalloc(o * o), *Dvv = zalloc(v * v), *Dov = zalloc(o * v), *Dvo = zalloc(v * o);
  double *Goooo = zalloc(o * o * o * o), *Gvvvv = zalloc(v * v * v * v), *Gooov = zalloc(o * o * o * v);
  double *Gvvvo = zalloc(v * v * v * o), *Govov = zalloc(o * v * o * v), *Goovv = zalloc(o * o * v * v);
#define T1(i, a) t1[O2(i, a)]
#define L1(i, a) l1[O2(i, a)]
#define T2(i, j, a, b) t2[T4(i, j, a, b)]
#define TAU(i, j, a, b) tau[T4(i, j, a, b)]
#define L2(i, j, a, b) l2[T4(i, j, a, b)]
  /* onepdm, ccdensity.cc:89-137 */
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++) {
      double d = s->tg_on ? s->tg_Doo[(size_t)i * no + j] : 0.0;
      for (int e = 0; e < nv; e++) d -= T1(i, e) * L1(j, e);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++)
          for (int f = 0; f < nv; f++) d -= T2(i, m, e, f) * L2(j, m, e, f);
      Doo[(size_t)i * no + j] = d;
    }
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++) {
      double d = s->tg_on ? s->tg_Dvv[(size_t)a * nv + b] : 0.0;
      for (int m = 0; m < no; m++) d += T1(m, b) * L1(m, a);
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++) d += T2(m, n, b, e) * L2(m, n, a, e);
      Dvv[(size_t)a * nv + b] = d;
    }
  for (int a = 0; a < nv; a++)
    for (int i = 0; i < no; i++) Dvo[(size_t)a * no + i] = L1(i, a);
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      double d = (s->tg_on ? s->tg_Dov[O2(i, a)] : 0.0) + 2.0 * T1(i, a);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) d += L1(m, e) * (2.0 * T2(i, m, a, e) - TAU(m, i, a, e));
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++)
              d -= L2(m, n, e, f) * (T1(m, a) * T2(i, n, e, f) + T1(i, e) * T2(m, n, a, f));
      Dov[O2(i, a)] = d;
    }
  double eone = 0.0;
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++) eone += F(i, j) * Doo[(size_t)i * no + j];
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++) eone += F(a + no, b + no) * Dvv[(size_t)a * nv + b];

  /* twopdm, ccdensity.cc:149-339 */
  double Eoooo = 0.0, Evvvv = 0.0, Eooov = 0.0, Evvvo = 0.0, Eovov = 0.0, Eoovv = 0.0;
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++)
        for (int l = 0; l < no; l++) {
          double g = 0.0;
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) g += TAU(i, j, e, f) * L2(k, l, e, f);
          Goooo[Q4(i, j, k, l, no, no, no)] = g;
          Eoooo += 0.5 * INT(i, j, k, l) * g;
        }
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int c = 0; c < nv; c++)
        for (int d = 0; d < nv; d++) {
          double g = 0.0;
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++) g += TAU(m, n, a, b) * L2(m, n, c, d);
          Gvvvv[Q4(a, b, c, d, nv, nv, nv)] = g;
          Evvvv += 0.5 * INT(a + no, b + no, c + no, d + no) * g;
        }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++)
        for (int a = 0; a < nv; a++) {
          double g = s->tg_on ? s->tg_Gooov[Q4(i, j, k, a, no, no, nv)] : 0.0;
          for (int e = 0; e < nv; e++) g -= L1(k, e) * (2.0 * TAU(i, j, e, a) - TAU(i, j, a, e));
          for (int e = 0; e < nv; e++) g -= T1(i, e) * L2(j, k, a, e);
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              for (int f = 0; f < nv; f++)
                g -= L2(k, m, e, f) * (2.0 * (T1(j, a) * T2(i, m, e, f) + T1(i, e) * T2(j, m, a, f)) -
                                       (T1(i, a) * T2(j, m, e, f) + T1(j, e) * T2(i, m, a, f)) -
                                       (T1(m, a) * T2(i, j, e, f) + T1(i, e) * T2(m, j, a, f) +
                                        T1(j, f) * T2(i, m, e, a)));
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              for (int f = 0; f < nv; f++) g += L2(k, m, e, f) * T1(m, a) * T1(i, e) * T1(j, f);
          Gooov[Q4(i, j, k, a, no, no, nv)] = g;
          Eooov += INT(i, j, k, a + no) * g;
        }
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int c = 0; c < nv; c++)
        for (int i = 0; i < no; i++) {
          double g = s->tg_on ? s->tg_Gvvvo[Q4(a, b, c, i, nv, nv, no)] : 0.0;
          for (int m = 0; m < no; m++) g += L1(m, c) * (2.0 * TAU(m, i, a, b) - TAU(i, m, a, b));
          for (int m = 0; m < no; m++) g += T1(m, a) * L2(i, m, b, c);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int e = 0; e < nv; e++)
                g += L2(n, m, c, e) * (2.0 * (T1(n, a) * T2(i, m, b, e) + T1(i, b) * T2(n, m, a, e)) -
                                       (T1(n, b) * T2(i, m, a, e) + T1(i, a) * T2(n, m, b, e)) -
                                       (T1(m, b) * T2(n, i, a, e) + T1(n, a) * T2(m, i, b, e) +
                                        T1(i, e) * T2(n, m, a, b)));
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int e = 0; e < nv; e++) g -= L2(n, m, c, e) * T1(m, b) * T1(n, a) * T1(i, e);
          Gvvvo[Q4(a, b, c, i, nv, nv, no)] = g;
          Evvvo += INT(a + no, b + no, c + no, i) * g;
        }
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++)
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++) {
          double g = -T1(i, a) * L1(j, b);
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++) g -= TAU(m, i, b, e) * L2(j, m, e, a) + T2(i, m, b, e) * L2(m, j, e, a);
          Govov[Q4(i, a, j, b, nv, no, nv)] = g;
          Eovov += INT(i, a + no, j, b + no) * g;
        }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double g = (s->tg_on ? s->tg_Goovv[T4(i, j, a, b)] : 0.0) + 4.0 * T1(i, a) * L1(j, b) +
                     4.0 * TAU(i, j, a, b) - 2.0 * TAU(i, j, b, a) + L2(i, j, a, b);
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              g += 2.0 * L1(m, e) *
                   (2.0 * T1(i, a) * (2.0 * T2(j, m, b, e) - T2(j, m, e, b)) -
                    T1(i, b) * (2.0 * T2(j, m, a, e) - T2(j, m, e, a)) -
                    T1(i, e) * (2.0 * TAU(j, m, b, a) - TAU(j, m, a, b)) -
                    T1(m, a) * (2.0 * T2(i, j, e, b) - T2(i, j, b, e)));
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              g += 4.0 * T2(i, m, a, e) * L2(m, j, e, b) - 2.0 * TAU(m, j, b, e) * L2(i, m, a, e);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int e = 0; e < nv; e++)
                for (int f = 0; f < nv; f++) {
                  const double l = L2(m, n, e, f);
                  g += 2.0 * l *
                       (0.5 * (T2(m, n, a, b) * T2(i, j, e, f) + T2(m, i, a, e) * T2(n, j, b, f) +
                               T2(n, j, a, e) * T2(i, m, f, b)) +
                        2.0 * (-TAU(i, j, a, e) * T2(m, n, b, f) - TAU(i, m, a, b) * T2(j, n, e, f) -
                               TAU(m, j, b, e) * T2(i, n, a, f) + T2(i, m, a, e) * T2(j, n, b, f)) -
                        (-TAU(i, j, b, e) * T2(m, n, a, f) - TAU(i, m, b, a) * T2(j, n, e, f) -
                         TAU(m, j, a, e) * T2(i, n, b, f) + T2(i, m, b, e) * T2(j, n, a, f)));
                  g += l * (T1(m, a) * T1(n, b) * T2(i, j, e, f) + T1(i, e) * T1(j, f) * T2(m, n, a, b) +
                            T1(m, a) * T1(i, e) * T2(n, j, b, f) + T1(n, b) * T1(j, f) * T2(m, i, a, e) +
                            T1(n, a) * T1(j, e) * T2(i, m, f, b) + T1(i, f) * T1(m, b) * T2(n, j, a, e));
                  g += l * T1(m, a) * T1(n, b) * T1(i, e) * T1(j, f);
                }
          Goovv[T4(i, j, a, b)] = g;
          Eoovv += 0.5 * INT(i, j, a + no, b + no) * g;
        }
#undef T1
#undef L1
#undef T2
#undef TAU
#undef L2
  if (E) {
    E[0] = eone;
    E[1] = Eoooo + Evvvv + Eooov + Evvvo + Eovov + Eoovv;
    E[2] = Eoooo; E[3] = Evvvv; E[4] = Eooov; E[5] = Evvvo; E[6] = Eovov; E[7] = Eoovv;
  }
#define OUT(dst, src, n) if (dst) memcpy(dst, src, (n) * sizeof(double))
  OUT(Doo_o, Doo, o * o); OUT(Dvv_o, Dvv, v * v); OUT(Dov_o, Dov, o * v); OUT(Dvo_o, Dvo, v * o);
  OUT(Goooo_o, Goooo, o * o * o * o); OUT(Gvvvv_o, Gvvvv, v * v * v * v); OUT(Gooov_o, Gooov, o * o * o * v);
  OUT(Gvvvo_o, Gvvvo, v * v * v * o); OUT(Govov_o, Govov, o * v * o * v); OUT(Goovv_o, Goovv, o * o * v * v);
#undef OUT
  free(Doo); free(Dvv); free(Dov); free(Dvo); free(Goooo); free(Gvvvv); free(Gooov); free(Gvvvo); free(Govov);
  free(Goovv);
}

/* ------------------------------------------------ (T) gradient terms
 * CCWfn::tgrad, ccwfn.cc:854-1098 (the in-core form, OOC=false, which
 * compute_energy runs for WFN=CCSD_T with DERTYPE=FIRST, :252-255), from the
 * current t1/t2 with t1s/t2s of build_tstar (:628-645):
 *   t3 = W3/D (the connected t3 of tcorr, :660-690);
 *   s1 (T3 --> Lambda1, :874-881);
 *   s2 = 4 X2 - 2 X2(ab swapped) with X2 = Y2 + Y2^T (T3 --> Lambda2,
 *        :884-915), then + Y2' + Y2'^T (Lambda3 --> Lambda2, :958-983);
 *   l3 = the six-permutation sum of l3_tmp (:917-956);
 *   the (T) density parts Doo, Dvv, Dov, Gooov, Gvvvo, Goovv (:985-1083).
 * Outputs (each may be NULL) in the layouts of orc_density; the terms are
 * kept in the state, and orc_tgrad_enable(1) makes orc_lambda add s1 and
 * s2/2 to the l1/l2 residuals (cclambda.cc:181, :243) and orc_density start
 * its blocks from the (T) parts (ccdensity.cc:43-74). */
void orc_tgrad(void *p, double *s1_o, double *s2_o, double *Doo_o, double *Dvv_o, double *Dov_o, double *Gooov_o,
               double *Gvvvo_o, double *Goovv_o) {
  orc_state *s = (orc_state *)p;
  const int no = s->no, nv = s->nv;
  const size_t o = no, v = nv, vvv = v * v * v, ooo = o * o * o;
  orc_build_tstar(s);
  const double *t1 = s->t1, *t2 = s->t2, *t1s = s->t1s, *t2s = s->t2s;
  double *t3 = zalloc(ooo * vvv), *l3t = zalloc(ooo * vvv), *l3 = zalloc(ooo * vvv);
#define X6(i, j, k, a, b, c) ((((size_t)(i) * no + (j)) * no + (k)) * vvv + V3(a, b, c))
#define T3(i, j, k, a, b, c) t3[X6(i, j, k, a, b, c)]
#define L3(i, j, k, a, b, c) l3[X6(i, j, k, a, b, c)]
#define T2(i, j, a, b) t2[T4(i, j, a, b)]
#define T2S(i, j, a, b) t2s[T4(i, j, a, b)]
#define DEN3(i, j, k, a, b, c) \
  (F(i, i) + F(j, j) + F(k, k) - F((a) + no, (a) + no) - F((b) + no, (b) + no) - F((c) + no, (c) + no))
#pragma omp parallel for collapse(3) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++)
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++) T3(i, j, k, a, b, c) = w3_elem(s, t2, i, j, k, a, b, c) / DEN3(i, j, k, a, b, c);
  free(s->tg_s1); free(s->tg_s2); free(s->tg_Doo); free(s->tg_Dvv); free(s->tg_Dov);
  free(s->tg_Gooov); free(s->tg_Gvvvo); free(s->tg_Goovv);
  double *s1 = s->tg_s1 = zalloc(o * v), *s2 = s->tg_s2 = zalloc(o * o * v * v);
  double *Doo = s->tg_Doo = zalloc(o * o), *Dvv = s->tg_Dvv = zalloc(v * v), *Dov = s->tg_Dov = zalloc(o * v);
  double *Gooov = s->tg_Gooov = zalloc(o * o * o * v), *Gvvvo = s->tg_Gvvvo = zalloc(v * v * v * o);
  double *Goovv = s->tg_Goovv = zalloc(o * o * v * v);
  double *Y2 = zalloc(o * o * v * v), *X2 = zalloc(o * o * v * v);

  /* T3 --> Lambda1, :874-881 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      double x = 0.0;
      for (int k = 0; k < no; k++)
        for (int l = 0; l < no; l++)
          for (int c = 0; c < nv; c++)
            for (int d = 0; d < nv; d++)
              x += 2.0 * (T3(i, k, l, a, c, d) - T3(k, i, l, a, c, d)) * LL(k, l, c + no, d + no);
      s1[O2(i, a)] += x;
    }
  /* T3 --> Lambda2, :884-915 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double x = 0.0;
          for (int k = 0; k < no; k++)
            for (int c = 0; c < nv; c++) x += (T3(i, j, k, a, b, c) - T3(k, j, i, a, b, c)) * F(k, c + no);
          for (int k = 0; k < no; k++)
            for (int c = 0; c < nv; c++)
              for (int l = 0; l < no; l++)
                x -= (2.0 * T3(j, k, l, b, a, c) - T3(j, l, k, b, a, c) - T3(l, k, j, b, a, c)) * INT(k, l, i, c + no);
          for (int k = 0; k < no; k++)
            for (int c = 0; c < nv; c++)
              for (int d = 0; d < nv; d++)
                x += (2.0 * T3(j, i, k, b, c, d) - T3(j, k, i, b, c, d) - T3(k, i, j, b, c, d)) *
                     INT(a + no, k, c + no, d + no);
          Y2[T4(i, j, a, b)] = x;
        }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) X2[T4(i, j, a, b)] += Y2[T4(i, j, a, b)] + Y2[T4(j, i, b, a)];
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) s2[T4(i, j, a, b)] += 4.0 * X2[T4(i, j, a, b)] - 2.0 * X2[T4(i, j, b, a)];

  /* Lambda3 amplitudes, :917-956 */
#pragma omp parallel for collapse(3) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++)
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++) {
              double x = F(i, a + no) * T2S(j, k, b, c) - F(j, a + no) * T2S(i, k, b, c);
              x += LL(i, j, a + no, b + no) * t1s[O2(k, c)] - LL(i, j, a + no, c + no) * t1s[O2(k, b)];
              for (int f = 0; f < nv; f++) {
                x += LL(f + no, j, a + no, b + no) * T2S(i, k, f, c);
                x -= INT(k, f + no, a + no, b + no) * T2S(i, j, c, f);
              }
              for (int n = 0; n < no; n++) {
                x -= LL(i, j, a + no, n) * T2S(n, k, b, c);
                x += INT(k, j, a + no, n) * T2S(n, i, b, c);
              }
              l3t[X6(i, j, k, a, b, c)] = x / DEN3(i, j, k, a, b, c);
            }
#pragma omp parallel for collapse(3) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++)
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++)
              L3(i, j, k, a, b, c) = l3t[X6(i, j, k, a, b, c)] + l3t[X6(i, k, j, a, c, b)] + l3t[X6(j, i, k, b, a, c)] +
                                     l3t[X6(j, k, i, b, c, a)] + l3t[X6(k, i, j, c, a, b)] + l3t[X6(k, j, i, c, b, a)];

  /* Lambda3 --> Lambda2, :958-983 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double y = 0.0;
          for (int k = 0; k < no; k++)
            for (int c = 0; c < nv; c++)
              for (int l = 0; l < no; l++) y -= L3(i, k, l, a, c, b) * INT(c + no, j, k, l);
          for (int k = 0; k < no; k++)
            for (int c = 0; c < nv; c++)
              for (int d = 0; d < nv; d++) y += L3(i, k, j, a, c, d) * INT(c + no, d + no, k, b + no);
          Y2[T4(i, j, a, b)] = y;
        }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) s2[T4(i, j, a, b)] += Y2[T4(i, j, a, b)] + Y2[T4(j, i, b, a)];

  /* (T) density contributions, :985-1083 */
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++) {
      double d = 0.0;
      for (int l = 0; l < no; l++)
        for (int m = 0; m < no; m++)
          for (int dd = 0; dd < nv; dd++)
            for (int e = 0; e < nv; e++)
              for (int f = 0; f < nv; f++) d -= 0.5 * T3(i, l, m, dd, e, f) * L3(j, l, m, dd, e, f);
      Doo[(size_t)i * no + j] = d;
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++) {
      double d = 0.0;
      for (int l = 0; l < no; l++)
        for (int m = 0; m < no; m++)
          for (int n = 0; n < no; n++)
            for (int dd = 0; dd < nv; dd++)
              for (int e = 0; e < nv; e++) d += 0.5 * T3(l, m, n, b, dd, e) * L3(l, m, n, a, dd, e);
      Dvv[(size_t)a * nv + b] = d;
    }
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      double d = 0.0;
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) d += (T3(m, n, i, e, f, a) - T3(m, i, n, e, f, a)) * T2S(m, n, e, f);
      Dov[O2(i, a)] = d;
    }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++)
        for (int a = 0; a < nv; a++) {
          double g = 0.0;
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              for (int f = 0; f < nv; f++)
                g -= (4.0 * T2(k, m, e, f) - 2.0 * T2(k, m, f, e)) *
                     (2.0 * T3(j, i, m, a, e, f) - T3(i, j, m, a, e, f) - T3(m, i, j, a, e, f));
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              for (int f = 0; f < nv; f++) g -= T2(k, m, e, f) * L3(m, j, i, f, a, e);
          Gooov[Q4(i, j, k, a, no, no, nv)] = g;
        }
#pragma omp parallel for collapse(2) schedule(static)
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int c = 0; c < nv; c++)
        for (int i = 0; i < no; i++) {
          double g = 0.0;
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int e = 0; e < nv; e++)
                g += (4.0 * T2(m, n, e, c) - 2.0 * T2(m, n, c, e)) *
                     (2.0 * T3(n, i, m, a, b, e) - T3(n, i, m, b, a, e) - T3(n, i, m, a, e, b));
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int e = 0; e < nv; e++) g += T2(m, n, e, c) * L3(n, i, m, a, b, e);
          Gvvvo[Q4(a, b, c, i, nv, nv, no)] = g;
        }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double g = 0.0;
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              g += 4.0 * t1[O2(m, e)] *
                   (2.0 * (T3(i, j, m, a, b, e) - T3(i, j, m, a, e, b)) - (T3(i, j, m, b, a, e) - T3(i, j, m, b, e, a)));
          Goovv[T4(i, j, a, b)] = g;
        }
#undef X6
#undef T3
#undef L3
#undef T2
#undef T2S
#undef DEN3
  free(t3); free(l3t); free(l3); free(Y2); free(X2);
#define OUT(dst, src, n) if (dst) memcpy(dst, src, (n) * sizeof(double))
  OUT(s1_o, s1, o * v); OUT(s2_o, s2, o * o * v * v); OUT(Doo_o, Doo, o * o); OUT(Dvv_o, Dvv, v * v);
  OUT(Dov_o, Dov, o * v); OUT(Gooov_o, Gooov, o * o * o * v); OUT(Gvvvo_o, Gvvvo, v * v * v * o);
  OUT(Goovv_o, Goovv, o * o * v * v);
#undef OUT
}

/* CCWfn::tgrad_ooc (ccwfn.cc:1150-1316, the OOC=true form): the same
 * outputs as orc_tgrad from W and V blocks formed one (i,j,k) at a time
 * (M3_ijk, N3_ijk, :1420-1515) and, for Doo, one (a,b,c) at a time
 * (M3_abc, N3_abc, :1456-1540), with no t3 or l3 array; Doo and Dvv get
 * their diagonals only.  The outputs replace those of orc_tgrad in the
 * state (orc_tgrad_enable applies to them alike). */
void orc_tgrad_ooc(void *p, double *s1_o, double *s2_o, double *Doo_o, double *Dvv_o, double *Dov_o,
                   double *Gooov_o, double *Gvvvo_o, double *Goovv_o) {
  orc_state *s = (orc_state *)p;
  const int no = s->no, nv = s->nv;
  const size_t o = no, v = nv, vvv = v * v * v, ooo = o * o * o;
  const double *t1 = s->t1, *t2 = s->t2;
  free(s->tg_s1); free(s->tg_s2); free(s->tg_Doo); free(s->tg_Dvv); free(s->tg_Dov);
  free(s->tg_Gooov); free(s->tg_Gvvvo); free(s->tg_Goovv);
  double *s1 = s->tg_s1 = zalloc(o * v), *s2 = s->tg_s2 = zalloc(o * o * v * v);
  double *Doo = s->tg_Doo = zalloc(o * o), *Dvv = s->tg_Dvv = zalloc(v * v), *Dov = s->tg_Dov = zalloc(o * v);
  double *Gooov = s->tg_Gooov = zalloc(o * o * o * v), *Gvvvo = s->tg_Gvvvo = zalloc(v * v * v * o);
  double *Goovv = s->tg_Goovv = zalloc(o * o * v * v);
  double *X2 = zalloc(o * o * v * v);
#define DEN3(i, j, k, a, b, c) \
  (F(i, i) + F(j, j) + F(k, k) - F((a) + no, (a) + no) - F((b) + no, (b) + no) - F((c) + no, (c) + no))
#define NUMN(i, j, k, a, b, c)                                                                              \
  (INT(i, j, (a) + no, (b) + no) * t1[O2(k, c)] + INT(i, k, (a) + no, (c) + no) * t1[O2(j, b)] +          \
   INT(j, k, (b) + no, (c) + no) * t1[O2(i, a)] + t2[T4(i, j, a, b)] * F(k, (c) + no) +                   \
   t2[T4(i, k, a, c)] * F(j, (b) + no) + t2[T4(j, k, b, c)] * F(i, (a) + no))
  double *M3 = zalloc(vvv), *N3 = zalloc(vvv), *X3 = zalloc(vvv), *Y3 = zalloc(vvv), *Z3 = zalloc(vvv);
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int k = 0; k < no; k++) {
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++) {
              const double den = DEN3(i, j, k, a, b, c);
              M3[V3(a, b, c)] = w3_elem(s, t2, i, j, k, a, b, c) / den;
              N3[V3(a, b, c)] = NUMN(i, j, k, a, b, c) / den;
            }
#define P6(X, a, b, c)                                                                                       \
  (8.0 * X[V3(a, b, c)] - 4.0 * X[V3(b, a, c)] - 4.0 * X[V3(a, c, b)] - 4.0 * X[V3(c, b, a)] +             \
   2.0 * X[V3(c, a, b)] + 2.0 * X[V3(b, c, a)])
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++) {
              X3[V3(a, b, c)] = P6(M3, a, b, c);
              Y3[V3(a, b, c)] = P6(N3, a, b, c);
              Z3[V3(a, b, c)] = 2.0 * (M3[V3(a, b, c)] - M3[V3(a, c, b)]) - (M3[V3(b, a, c)] - M3[V3(b, c, a)]);
            }
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++)
            for (int c = 0; c < nv; c++) {
              const double m = M3[V3(a, b, c)], xy = 2.0 * X3[V3(a, b, c)] + Y3[V3(a, b, c)];
              Dvv[(size_t)a * v + a] += 0.5 * m * (X3[V3(a, b, c)] + Y3[V3(a, b, c)]);
              Dov[O2(i, a)] += (m - M3[V3(c, b, a)]) * (4.0 * t2[T4(j, k, b, c)] - 2.0 * t2[T4(j, k, c, b)]);
              s1[O2(i, a)] += (4.0 * m - 4.0 * M3[V3(b, a, c)] - 2.0 * M3[V3(a, c, b)] + 2.0 * M3[V3(c, a, b)]) *
                              INT(j, k, b + no, c + no);
              Goovv[T4(i, j, a, b)] += 4.0 * t1[O2(k, c)] * Z3[V3(a, b, c)];
              for (int l = 0; l < no; l++) {
                X2[T4(i, l, a, b)] -= xy * INT(j, k, l, c + no);
                Gooov[(((size_t)j * o + i) * o + l) * v + a] -= xy * t2[T4(l, k, b, c)];
              }
              for (int d = 0; d < nv; d++) {
                X2[T4(i, j, a, d)] += xy * INT(d + no, k, b + no, c + no);
                Gvvvo[(((size_t)a * v + b) * v + d) * o + j] += xy * t2[T4(k, i, c, d)];
              }
            }
      }
  free(M3); free(N3); free(X3); free(Y3); free(Z3);
  M3 = zalloc(ooo); N3 = zalloc(ooo); X3 = zalloc(ooo); Y3 = zalloc(ooo);
#define O3(i, j, k) ((((size_t)(i)) * no + (j)) * no + (k))
#define Q6(X, i, j, k)                                                                                       \
  (8.0 * X[O3(i, j, k)] - 4.0 * X[O3(j, i, k)] - 4.0 * X[O3(i, k, j)] - 4.0 * X[O3(k, j, i)] +             \
   2.0 * X[O3(k, i, j)] + 2.0 * X[O3(j, k, i)])
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int c = 0; c < nv; c++) {
        for (int i = 0; i < no; i++)
          for (int j = 0; j < no; j++)
            for (int k = 0; k < no; k++) {
              const double den = DEN3(i, j, k, a, b, c);
              M3[O3(i, j, k)] = w3_elem(s, t2, i, j, k, a, b, c) / den;
              N3[O3(i, j, k)] = NUMN(i, j, k, a, b, c) / den;
            }
        for (int i = 0; i < no; i++)
          for (int j = 0; j < no; j++)
            for (int k = 0; k < no; k++) {
              X3[O3(i, j, k)] = Q6(M3, i, j, k);
              Y3[O3(i, j, k)] = Q6(N3, i, j, k);
            }
        for (int i = 0; i < no; i++)
          for (int j = 0; j < no; j++)
            for (int k = 0; k < no; k++)
              Doo[(size_t)i * o + i] -= 0.5 * M3[O3(i, j, k)] * (X3[O3(i, j, k)] + Y3[O3(i, j, k)]);
      }
  free(M3); free(N3); free(X3); free(Y3);
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) s2[T4(i, j, a, b)] = X2[T4(i, j, a, b)] + X2[T4(j, i, b, a)];
  free(X2);
#undef DEN3
#undef NUMN
#undef P6
#undef Q6
#undef O3
#define OUT(dst, src, n) if (dst) memcpy(dst, src, (n) * sizeof(double))
  OUT(s1_o, s1, o * v); OUT(s2_o, s2, o * o * v * v); OUT(Doo_o, Doo, o * o); OUT(Dvv_o, Dvv, v * v);
  OUT(Dov_o, Dov, o * v); OUT(Gooov_o, Gooov, o * o * o * v); OUT(Gvvvo_o, Gvvvo, v * v * v * o);
  OUT(Goovv_o, Goovv, o * o * v * v);
#undef OUT
}

/* 1: orc_lambda and orc_density include the terms of the last orc_tgrad */
void orc_tgrad_enable(void *p, int on) {
  orc_state *s = (orc_state *)p;
  s->tg_on = on && s->tg_s1;
}

/* -------------------------------------------------- linear response
 * CCPert (ccpert.cc) and CCResp::linresp (ccresp.cc:27-124): the perturbed
 * amplitudes of a one-electron perturbation A (an nact x nact MO matrix,
 * Perturbation::prop_p, perturbation.cc) at frequency omega over the HBAR of
 * orc_lambda_init and the Lambda amplitudes of orc_lambda, and the linear
 * response function of two of them. */
#define HB(k) (s->hb[k])
typedef struct {
  orc_state *s;
  double omega;
  double *pert;                                        /* nact x nact */
  double *Aov, *Aoo, *Avv, *Avo, *Aovoo, *Avvvo, *Avvoo; /* pertbar, ccpert.cc:112-176 */
  double *D1, *D2;
  double *X1, *X1o, *X2, *X2o, *Y1, *Y1o, *Y2, *Y2o;
  double *Goo, *Gvv;
} orc_pert;

#define PA(p, q) (P->pert[(size_t)(p) * NA + (q)])
#define AOV(m, e) P->Aov[O2(m, e)]
#define AOO(m, i) P->Aoo[(size_t)(m) * no + (i)]
#define AVV(a, e) P->Avv[(size_t)(a) * nv + (e)]
#define AVO(a, i) P->Avo[(size_t)(a) * no + (i)]
#define AOVOO(m, b, i, j) P->Aovoo[Q4(m, b, i, j, nv, no, no)]
#define AVVVO(a, b, e, i) P->Avvvo[Q4(a, b, e, i, nv, nv, no)]
#define AVVOO(a, b, i, j) P->Avvoo[Q4(a, b, i, j, nv, no, no)]

/* CCPert::pertbar, ccpert.cc:112-176: the similarity-transformed perturbation */
static void pert_bar(orc_pert *P) {
  orc_state *s = P->s;
  const int no = s->no, nv = s->nv;
  const double *t1 = s->t1, *t2 = s->t2;
  for (int m = 0; m < no; m++)
    for (int e = 0; e < nv; e++) AOV(m, e) = PA(m, e + no);
  for (int m = 0; m < no; m++)
    for (int i = 0; i < no; i++) {
      double x = PA(m, i);
      for (int e = 0; e < nv; e++) x += t1[O2(i, e)] * PA(m, e + no);
      AOO(m, i) = x;
    }
  for (int a = 0; a < nv; a++)
    for (int e = 0; e < nv; e++) {
      double x = PA(a + no, e + no);
      for (int m = 0; m < no; m++) x -= t1[O2(m, a)] * PA(m, e + no);
      AVV(a, e) = x;
    }
  for (int a = 0; a < nv; a++)
    for (int i = 0; i < no; i++) {
      double x = PA(a + no, i);
      for (int e = 0; e < nv; e++) x += t1[O2(i, e)] * PA(a + no, e + no);
      for (int m = 0; m < no; m++) x -= t1[O2(m, a)] * PA(m, i);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++)
          x += (2.0 * t2[T4(m, i, e, a)] - t2[T4(i, m, e, a)] - t1[O2(i, e)] * t1[O2(m, a)]) * PA(m, e + no);
      AVO(a, i) = x;
    }
  for (int m = 0; m < no; m++)
    for (int b = 0; b < nv; b++)
      for (int i = 0; i < no; i++)
        for (int j = 0; j < no; j++) {
          double x = 0.0;
          for (int e = 0; e < nv; e++) x += t2[T4(i, j, e, b)] * PA(m, e + no);
          AOVOO(m, b, i, j) = x;
        }
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int e = 0; e < nv; e++)
        for (int i = 0; i < no; i++) {
          double x = 0.0;
          for (int m = 0; m < no; m++) x -= t2[T4(m, i, a, b)] * PA(m, e + no);
          AVVVO(a, b, e, i) = x;
        }
  for (int a = 0; a < nv; a++)
    for (int b = 0; b < nv; b++)
      for (int i = 0; i < no; i++)
        for (int j = 0; j < no; j++) {
          double x = 0.0;
          for (int e = 0; e < nv; e++) x += t2[T4(i, j, e, b)] * AVV(a, e);
          for (int m = 0; m < no; m++) x -= t2[T4(m, j, a, b)] * AOO(m, i);
          AVVOO(a, b, i, j) = x;
        }
}

/* CCPert::CCPert, ccpert.cc:11-75: pertbar, the HBAR-diagonal denominators
 * and the guesses X = Y = Abar / (D + omega).  Needs orc_lambda_init (HBAR)
 * and the Lambda amplitudes (orc_lambda or orc_set_lambda). */
void *orc_pert_create(void *p, const double *pert, double omega) {
  orc_state *s = (orc_state *)p;
  const int no = s->no, nv = s->nv;
  const size_t o = no, v = nv, n = s->nact, ov = o * v;
  orc_pert *P = (orc_pert *)calloc(1, sizeof(orc_pert));
  P->s = s; P->omega = omega;
  P->pert = zalloc(n * n); memcpy(P->pert, pert, n * n * sizeof(double));
  P->Aov = zalloc(ov); P->Aoo = zalloc(o * o); P->Avv = zalloc(v * v); P->Avo = zalloc(ov);
  P->Aovoo = zalloc(o * v * o * o); P->Avvvo = zalloc(v * v * v * o); P->Avvoo = zalloc(v * v * o * o);
  P->D1 = zalloc(ov); P->D2 = zalloc(ov * ov);
  P->X1 = zalloc(ov); P->X1o = zalloc(ov); P->Y1 = zalloc(ov); P->Y1o = zalloc(ov);
  P->X2 = zalloc(ov * ov); P->X2o = zalloc(ov * ov); P->Y2 = zalloc(ov * ov); P->Y2o = zalloc(ov * ov);
  P->Goo = zalloc(o * o); P->Gvv = zalloc(v * v);
  pert_bar(P);
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      P->D1[O2(i, a)] = Hoo_(i, i) - Hvv_(a, a);
      P->X1[O2(i, a)] = P->Y1[O2(i, a)] = AVO(a, i) / (P->D1[O2(i, a)] + omega);
    }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          const double d = Hoo_(i, i) + Hoo_(j, j) - Hvv_(a, a) - Hvv_(b, b);
          P->D2[T4(i, j, a, b)] = d;
          P->X2[T4(i, j, a, b)] = P->Y2[T4(i, j, a, b)] = (AVVOO(a, b, i, j) + AVVOO(b, a, j, i)) / (d + omega);
        }
  return P;
}

void orc_pert_destroy(void *q) {
  orc_pert *P = (orc_pert *)q;
  if (!P) return;
  free(P->pert); free(P->Aov); free(P->Aoo); free(P->Avv); free(P->Avo); free(P->Aovoo); free(P->Avvvo);
  free(P->Avvoo); free(P->D1); free(P->D2); free(P->X1); free(P->X1o); free(P->X2); free(P->X2o);
  free(P->Y1); free(P->Y1o); free(P->Y2); free(P->Y2o); free(P->Goo); free(P->Gvv);
  free(P);
}

/* CCPert::build_X1, ccpert.cc:382-407: raw residual from (X1o, X2o) into R1 */
static void pert_X1(orc_pert *P, const double *X1o, const double *X2o, double *R1) {
  orc_state *s = P->s;
  const int no = s->no, nv = s->nv;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      double x = AVO(a, i) - P->omega * X1o[O2(i, a)];
      for (int e = 0; e < nv; e++) x += X1o[O2(i, e)] * Hvv_(a, e);
      for (int m = 0; m < no; m++) x -= X1o[O2(m, a)] * Hoo_(m, i);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) x += X1o[O2(m, e)] * (2.0 * Hovvo_(m, a, e, i) - Hovov_(m, a, i, e));
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) x += Hov_(m, e) * (2.0 * X2o[T4(m, i, e, a)] - X2o[T4(i, m, e, a)]);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++)
          for (int f = 0; f < nv; f++) x += X2o[T4(i, m, e, f)] * (2.0 * Hvovv_(a, m, e, f) - Hvovv_(a, m, f, e));
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++) x -= X2o[T4(m, n, a, e)] * (2.0 * Hooov_(m, n, i, e) - Hooov_(n, m, i, e));
      R1[O2(i, a)] = x;
    }
}

/* CCPert::build_X2, ccpert.cc:409-470: R2 = Z + Z^T from (X1o, X2o) */
static void pert_X2(orc_pert *P, const double *X1o, const double *X2o, double *R2) {
  orc_state *s = P->s;
  const int no = s->no, nv = s->nv;
  const double *t2 = s->t2;
  double *Zvv = zalloc((size_t)nv * nv), *Zoo = zalloc((size_t)no * no), *Z = zalloc((size_t)no * no * nv * nv);
  for (int a = 0; a < nv; a++)
    for (int e = 0; e < nv; e++) {
      double x = 0.0;
      for (int m = 0; m < no; m++)
        for (int f = 0; f < nv; f++) x += (2.0 * Hvovv_(a, m, e, f) - Hvovv_(a, m, f, e)) * X1o[O2(m, f)];
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int f = 0; f < nv; f++) x -= LL(m, n, e + no, f + no) * X2o[T4(m, n, a, f)];
      Zvv[(size_t)a * nv + e] = x;
    }
  for (int m = 0; m < no; m++)
    for (int i = 0; i < no; i++) {
      double x = 0.0;
      for (int n = 0; n < no; n++)
        for (int e = 0; e < nv; e++) x -= (2.0 * Hooov_(m, n, i, e) - Hooov_(n, m, i, e)) * X1o[O2(n, e)];
      for (int n = 0; n < no; n++)
        for (int e = 0; e < nv; e++)
          for (int f = 0; f < nv; f++) x -= LL(m, n, e + no, f + no) * X2o[T4(i, n, e, f)];
      Zoo[(size_t)m * no + i] = x;
    }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double x = AVVOO(a, b, i, j) - 0.5 * P->omega * X2o[T4(i, j, a, b)];
          for (int e = 0; e < nv; e++) x += X1o[O2(i, e)] * Hvvvo_(a, b, e, j);
          for (int m = 0; m < no; m++) x -= X1o[O2(m, a)] * Hovoo_(m, b, i, j);
          for (int m = 0; m < no; m++) x += Zoo[(size_t)m * no + i] * t2[T4(m, j, a, b)];
          for (int e = 0; e < nv; e++) x += Zvv[(size_t)a * nv + e] * t2[T4(i, j, e, b)];
          for (int e = 0; e < nv; e++) x += X2o[T4(i, j, e, b)] * Hvv_(a, e);
          for (int m = 0; m < no; m++) x -= X2o[T4(m, j, a, b)] * Hoo_(m, i);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++) x += 0.5 * X2o[T4(m, n, a, b)] * Hoooo_(m, n, i, j);
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) x += 0.5 * X2o[T4(i, j, e, f)] * Hvvvv_(a, b, e, f);
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              x -= X2o[T4(i, m, e, b)] * Hovov_(m, a, j, e) + X2o[T4(i, m, e, a)] * Hovvo_(m, b, e, j);
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++)
              x += X2o[T4(m, i, e, a)] * (2.0 * Hovvo_(m, b, e, j) - Hovov_(m, b, j, e));
          Z[T4(i, j, a, b)] = x;
        }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) R2[T4(i, j, a, b)] = Z[T4(i, j, a, b)] + Z[T4(j, i, b, a)];
  free(Zvv); free(Zoo); free(Z);
}

/* CCPert::build_G, ccpert.cc:489-513, from Y2o */
static void pert_G(orc_pert *P, const double *Y2o) {
  orc_state *s = P->s;
  const int no = s->no, nv = s->nv;
  const double *t2 = s->t2;
  for (int m = 0; m < no; m++)
    for (int i = 0; i < no; i++) {
      double x = 0.0;
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++)
          for (int j = 0; j < no; j++) x += t2[T4(m, j, a, b)] * Y2o[T4(i, j, a, b)];
      P->Goo[(size_t)m * no + i] = x;
    }
  for (int a = 0; a < nv; a++)
    for (int e = 0; e < nv; e++) {
      double x = 0.0;
      for (int i = 0; i < no; i++)
        for (int j = 0; j < no; j++)
          for (int b = 0; b < nv; b++) x -= t2[T4(i, j, e, b)] * Y2o[T4(i, j, a, b)];
      P->Gvv[(size_t)a * nv + e] = x;
    }
}

/* CCPert::build_Y1, ccpert.cc:516-703: the left-hand residual from (Y1o,
 * Y2o) and G, with the terms of the right-hand amplitudes X1/X2, the Lambda
 * amplitudes and the perturbation (the reference's (i) .. (t)) */
static void pert_Y1(orc_pert *P, const double *Y1o, const double *Y2o, double *R1) {
  orc_state *s = P->s;
  const int no = s->no, nv = s->nv;
  const double *t2 = s->t2, *l1 = s->l1, *l2 = s->l2, *X1 = P->X1, *X2 = P->X2;
  const double w = P->omega;
#define Y1_(i, a) Y1o[O2(i, a)]
#define Y2_(i, j, a, b) Y2o[T4(i, j, a, b)]
#define X1_(i, a) X1[O2(i, a)]
#define X2_(i, j, a, b) X2[T4(i, j, a, b)]
#define l1_(i, a) l1[O2(i, a)]
#define l2_(i, j, a, b) l2[T4(i, j, a, b)]
#define T2_(i, j, a, b) t2[T4(i, j, a, b)]
#define GVV(a, e) P->Gvv[(size_t)(a) * nv + (e)]
#define GOO(m, i) P->Goo[(size_t)(m) * no + (i)]
#define LV(p, q, r, t) LL(p, q, (r) + no, (t) + no)
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      double x = 2 * PA(i, a + no) + w * Y1_(i, a);
      for (int e = 0; e < nv; e++) x += Y1_(i, e) * Hvv_(e, a);
      for (int m = 0; m < no; m++) x -= Y1_(m, a) * Hoo_(i, m);
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) x += Y1_(m, e) * (2 * Hovvo_(i, e, a, m) - Hovov_(i, e, m, a));
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++)
          for (int f = 0; f < nv; f++) x += Y2_(i, m, e, f) * Hvvvo_(e, f, a, m);
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++) x -= Y2_(m, n, a, e) * Hovoo_(i, e, m, n);
      for (int e = 0; e < nv; e++)
        for (int f = 0; f < nv; f++) x -= GVV(e, f) * (2 * Hvovv_(e, i, f, a) - Hvovv_(e, i, a, f));
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++) x -= GOO(m, n) * (2 * Hooov_(m, i, n, a) - Hooov_(i, m, n, a));
      /* (i) */
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) x += 2.0 * X1_(m, e) * LV(i, m, a, e);
      /* (m) */
      for (int m = 0; m < no; m++) x -= l1_(m, a) * AOO(i, m);
      for (int e = 0; e < nv; e++) x += l1_(i, e) * AVV(e, a);
      /* (n) */
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++)
          for (int f = 0; f < nv; f++) x += l2_(i, m, f, e) * AVVVO(f, e, a, m);
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++) {
            x -= 0.5 * l2_(m, n, e, a) * AOVOO(i, e, n, m);
            x -= 0.5 * l2_(m, n, a, e) * AOVOO(i, e, m, n);
          }
      /* (q) */
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) {
          const double xm = X1_(m, e);
          x -= xm * Hov_(m, a) * l1_(i, e);
          x -= xm * Hov_(i, e) * l1_(m, a);
          for (int n = 0; n < no; n++) {
            x -= xm * (2 * Hooov_(m, i, n, a) - Hooov_(i, m, n, a)) * l1_(n, e);
            x -= xm * (2 * Hooov_(i, m, n, e) - Hooov_(m, i, n, e)) * l1_(n, a);
          }
          for (int f = 0; f < nv; f++) {
            x += xm * (2 * Hvovv_(f, m, a, e) - Hvovv_(f, m, e, a)) * l1_(i, f);
            x += xm * (2 * Hvovv_(f, i, e, a) - Hvovv_(f, i, a, e)) * l1_(m, f);
          }
        }
      /* (r) */
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) {
              const double xm = X2_(m, n, e, f);
              x += 2.0 * xm * LV(i, m, a, e) * l1_(n, f);
              x -= xm * LV(i, m, a, f) * l1_(n, e);
              x -= xm * LV(m, i, e, f) * l1_(n, a);
              x -= xm * LV(n, m, f, a) * l1_(i, e);
            }
      /* (s) */
      for (int m = 0; m < no; m++)
        for (int e = 0; e < nv; e++) {
          const double xm = X1_(m, e);
          for (int n = 0; n < no; n++)
            for (int f = 0; f < nv; f++) {
              x -= xm * Hovov_(m, f, n, a) * l2_(n, i, e, f);
              x -= xm * Hovov_(i, f, n, e) * l2_(n, m, a, f);
              x -= xm * Hovvo_(m, f, a, n) * l2_(n, i, f, e);
              x -= xm * Hovvo_(i, f, e, n) * l2_(n, m, f, a);
            }
          for (int f = 0; f < nv; f++)
            for (int g = 0; g < nv; g++) {
              x += 0.5 * xm * Hvvvv_(f, g, a, e) * l2_(i, m, f, g);
              x += 0.5 * xm * Hvvvv_(f, g, e, a) * l2_(m, i, f, g);
            }
          for (int n = 0; n < no; n++)
            for (int q = 0; q < no; q++) {
              x += 0.5 * xm * Hoooo_(i, m, n, q) * l2_(n, q, a, e);
              x += 0.5 * xm * Hoooo_(m, i, n, q) * l2_(n, q, e, a);
            }
        }
      /* the four-factor terms, ccpert.cc:629-640 */
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++)
              for (int j = 0; j < no; j++)
                for (int b = 0; b < nv; b++) {
                  x -= LV(i, n, a, f) * X1_(n, b) * T2_(j, m, f, e) * l2_(j, m, b, e);
                  x -= LV(m, i, e, f) * X1_(m, e) * T2_(j, n, f, b) * l2_(j, n, a, b);
                  x -= LV(m, n, e, a) * X1_(m, e) * T2_(j, n, f, b) * l2_(j, i, f, b);
                  x -= LV(i, n, a, f) * X1_(j, f) * T2_(n, m, e, b) * l2_(j, m, e, b);
                }
      /* (t) */
      for (int m = 0; m < no; m++)
        for (int n = 0; n < no; n++)
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) {
              const double xm = X2_(m, n, e, f);
              x -= xm * Hov_(m, a) * l2_(n, i, f, e);
              x -= xm * Hov_(i, e) * l2_(n, m, f, a);
              for (int g = 0; g < nv; g++) {
                x -= xm * Hvovv_(g, n, e, a) * l2_(i, m, f, g);
                x -= xm * Hvovv_(g, n, a, e) * l2_(m, i, f, g);
                x -= xm * Hvovv_(g, i, e, f) * l2_(n, m, a, g);
                x += xm * (2 * Hvovv_(g, m, a, e) - Hvovv_(g, m, e, a)) * l2_(n, i, f, g);
                x += xm * (2 * Hvovv_(g, i, e, a) - Hvovv_(g, i, a, e)) * l2_(n, m, f, g);
              }
              for (int q = 0; q < no; q++) {
                x += xm * Hooov_(m, n, q, a) * l2_(q, i, e, f);
                x += xm * Hooov_(i, n, q, e) * l2_(q, m, a, f);
                x += xm * Hooov_(m, i, q, f) * l2_(q, n, e, a);
                x -= xm * (2 * Hooov_(m, i, q, a) - Hooov_(i, m, q, a)) * l2_(n, q, f, e);
                x -= xm * (2 * Hooov_(i, m, q, e) - Hooov_(m, i, q, e)) * l2_(n, q, f, a);
              }
            }
      R1[O2(i, a)] = x;
    }
}

/* CCPert::build_Y2, ccpert.cc:705-882: R2 = Z + Z^T from (Y1o, Y2o) and G,
 * with the terms (o) .. (x) */
static void pert_Y2(orc_pert *P, const double *Y1o, const double *Y2o, double *R2) {
  orc_state *s = P->s;
  const int no = s->no, nv = s->nv;
  const double *l1 = s->l1, *l2 = s->l2, *X1 = P->X1, *X2 = P->X2;
  const double w = P->omega;
  double *Z = zalloc((size_t)no * no * nv * nv);
#define IV(p, q, r, t) INT(p, q, (r) + no, (t) + no)
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) {
          double x = 0.5 * w * Y2_(i, j, a, b);
          x += 2.0 * Y1_(i, a) * Hov_(j, b) - Y1_(j, a) * Hov_(i, b);
          for (int e = 0; e < nv; e++) x += Y2_(i, j, e, b) * Hvv_(e, a);
          for (int m = 0; m < no; m++) x -= Y2_(m, j, a, b) * Hoo_(i, m);
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++) x += 0.5 * Y2_(m, n, a, b) * Hoooo_(i, j, m, n);
          for (int e = 0; e < nv; e++)
            for (int f = 0; f < nv; f++) x += 0.5 * Y2_(i, j, e, f) * Hvvvv_(e, f, a, b);
          for (int e = 0; e < nv; e++) x += Y1_(i, e) * (2 * Hvovv_(e, j, a, b) - Hvovv_(e, j, b, a));
          for (int m = 0; m < no; m++) x -= Y1_(m, b) * (2 * Hooov_(j, i, m, a) - Hooov_(i, j, m, a));
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++) {
              x += Y2_(m, j, e, b) * (2 * Hovvo_(i, e, a, m) - Hovov_(i, e, m, a));
              x -= Y2_(m, i, b, e) * Hovov_(j, e, m, a);
              x -= Y2_(m, i, e, b) * Hovvo_(j, e, a, m);
            }
          for (int e = 0; e < nv; e++) x += GVV(a, e) * LV(i, j, e, b);
          for (int m = 0; m < no; m++) x -= GOO(m, i) * LV(m, j, a, b);
          /* (o), (p) */
          x += 2 * l1_(j, b) * PA(i, a + no) - l1_(i, b) * PA(j, a + no);
          for (int e = 0; e < nv; e++) x += l2_(i, j, e, b) * AVV(e, a);
          for (int m = 0; m < no; m++) x -= l2_(m, j, a, b) * AOO(i, m);
          /* (u) */
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++) {
              const double xm = X1_(m, e);
              x -= xm * l1_(j, a) * LV(m, i, e, b);
              x -= xm * l1_(m, b) * LV(i, j, a, e);
              x -= xm * l1_(i, e) * LV(j, m, b, a);
              x += 2.0 * xm * l1_(j, b) * LV(i, m, a, e);
            }
          /* (w) */
          for (int m = 0; m < no; m++)
            for (int e = 0; e < nv; e++) {
              const double xm = X1_(m, e);
              x -= Hov_(m, a) * l2_(j, i, b, e) * xm;
              x -= Hov_(i, e) * l2_(j, m, b, a) * xm;
              for (int f = 0; f < nv; f++) {
                x -= Hvovv_(f, m, b, a) * l2_(j, i, f, e) * xm;
                x -= Hvovv_(f, j, e, a) * l2_(m, i, f, b) * xm;
                x -= Hvovv_(f, i, b, e) * l2_(j, m, f, a) * xm;
                x += (2 * Hvovv_(f, m, a, e) - Hvovv_(f, m, e, a)) * l2_(j, i, b, f) * xm;
                x += (2 * Hvovv_(f, i, e, a) - Hvovv_(f, i, a, e)) * l2_(j, m, b, f) * xm;
              }
              for (int n = 0; n < no; n++) {
                x += Hooov_(j, m, n, a) * l2_(n, i, b, e) * xm;
                x += Hooov_(m, j, n, a) * l2_(n, i, e, b) * xm;
                x += Hooov_(j, i, n, e) * l2_(n, m, b, a) * xm;
                x -= (2 * Hooov_(m, i, n, a) - Hooov_(i, m, n, a)) * l2_(j, n, b, e) * xm;
                x -= (2 * Hooov_(i, m, n, e) - Hooov_(m, i, n, e)) * l2_(j, n, b, a) * xm;
              }
            }
          /* (x) */
          for (int m = 0; m < no; m++)
            for (int n = 0; n < no; n++)
              for (int e = 0; e < nv; e++)
                for (int f = 0; f < nv; f++) {
                  const double xm = X2_(m, n, e, f);
                  x += 0.25 * IV(n, m, a, b) * l2_(j, i, e, f) * xm;
                  x += 0.25 * IV(m, n, a, b) * l2_(j, i, f, e) * xm;
                  x += 0.25 * IV(j, n, a, e) * l2_(m, i, f, b) * xm;
                  x += 0.25 * IV(j, m, a, f) * l2_(n, i, e, b) * xm;
                  x += 0.25 * IV(n, j, a, e) * l2_(m, i, b, f) * xm;
                  x += 0.25 * IV(m, j, a, f) * l2_(n, i, b, e) * xm;
                  x += 0.25 * IV(j, n, e, a) * l2_(i, m, f, b) * xm;
                  x += 0.25 * IV(j, m, f, a) * l2_(i, n, e, b) * xm;
                  x += 0.25 * IV(n, j, e, a) * l2_(i, m, b, f) * xm;
                  x += 0.25 * IV(m, j, f, a) * l2_(i, n, b, e) * xm;
                  x += 0.25 * IV(j, i, e, f) * l2_(n, m, a, b) * xm;
                  x += 0.25 * IV(j, i, f, e) * l2_(m, n, a, b) * xm;
                  x -= 0.5 * LV(i, m, a, f) * l2_(j, n, b, e) * xm;
                  x -= 0.5 * LV(i, n, a, e) * l2_(j, m, b, f) * xm;
                  x -= 0.5 * LV(m, i, e, f) * l2_(j, n, b, a) * xm;
                  x -= 0.5 * LV(m, n, e, a) * l2_(j, i, b, f) * xm;
                  x -= 0.5 * LV(n, m, f, a) * l2_(j, i, b, e) * xm;
                  x -= 0.5 * LV(n, i, f, e) * l2_(j, m, b, a) * xm;
                  x -= LV(i, j, a, e) * l2_(n, m, f, b) * xm;
                  x -= LV(i, m, a, b) * l2_(n, j, f, e) * xm;
                  x -= LV(m, j, e, a) * l2_(n, i, f, b) * xm;
                  x += 2.0 * LV(i, m, a, e) * l2_(n, j, f, b) * xm;
                }
          Z[T4(i, j, a, b)] = x;
        }
  for (int i = 0; i < no; i++)
    for (int j = 0; j < no; j++)
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++) R2[T4(i, j, a, b)] = Z[T4(i, j, a, b)] + Z[T4(j, i, b, a)];
  free(Z);
#undef IV
}

/* CCPert::pseudoresponse, ccpert.cc:229-249 */
static double pert_pseudo(orc_pert *P, const double *X1, const double *X2) {
  orc_state *s = P->s;
  const int no = s->no, nv = s->nv;
  double p1 = 0.0, p2 = 0.0;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      p1 += 2.0 * X1[O2(i, a)] * AVO(a, i);
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++)
          p2 += 2.0 * (2.0 * X2[T4(i, j, a, b)] - X2[T4(i, j, b, a)]) * AVVOO(a, b, i, j);
    }
  return -2.0 * (p1 + p2);
}

double orc_pert_pseudoresponse(void *q, int hand) {
  orc_pert *P = (orc_pert *)q;
  return hand == 0 ? pert_pseudo(P, P->X1, P->X2) : pert_pseudo(P, P->Y1, P->Y2);
}

/* One residual evaluation at the given amplitudes (the object's own are not
 * touched, except that hand 1 uses its X1/X2): hand 0 build_X1/build_X2,
 * hand 1 build_G/build_Y1/build_Y2; raw R1, R2 (before the denominators). */
void orc_pert_residual(void *q, int hand, const double *A1, const double *A2, double *R1, double *R2) {
  orc_pert *P = (orc_pert *)q;
  if (hand == 0) {
    pert_X1(P, A1, A2, R1);
    pert_X2(P, A1, A2, R2);
  } else {
    pert_G(P, A2);
    pert_Y1(P, A1, A2, R1);
    pert_Y2(P, A1, A2, R2);
  }
}

/* CCPert::solve, ccpert.cc:178-225: amp_save, build, increment_amps
 * (x = x_old + R/(D + omega), rms the 2-norm of the raw R), break on rms <
 * conv before DIIS, else DIIS (the Pulay step of orc_ccsd, extrapolating from
 * min_vecs stored vectors).  hand 0: X (right), 1: Y (left; needs the
 * converged X).  out: pseudoresponse, rms, niter.  Returns 0 on convergence. */
int orc_pert_solve(void *q, int hand, int maxiter, double conv, int do_diis, int min_vecs, double *out) {
  orc_pert *P = (orc_pert *)q;
  orc_state *s = P->s;
  const size_t ov = (size_t)s->no * s->nv;
  double **A1 = hand == 0 ? &P->X1 : &P->Y1, **A1o = hand == 0 ? &P->X1o : &P->Y1o;
  double **A2 = hand == 0 ? &P->X2 : &P->Y2, **A2o = hand == 0 ? &P->X2o : &P->Y2o;
  diis_free(s);
  double rms = 0.0;
  int niter = 0;
  for (int iter = 1; iter <= maxiter; iter++) {
    double *x = *A1; *A1 = *A1o; *A1o = x;
    x = *A2; *A2 = *A2o; *A2o = x;
    orc_pert_residual(P, hand, *A1o, *A2o, *A1, *A2);
    double r1 = 0.0, r2 = 0.0;
    for (size_t k = 0; k < ov; k++) {
      r1 += (*A1)[k] * (*A1)[k];
      (*A1)[k] = (*A1o)[k] + (*A1)[k] / (P->D1[k] + P->omega);
    }
    for (size_t k = 0; k < ov * ov; k++) {
      r2 += (*A2)[k] * (*A2)[k];
      (*A2)[k] = (*A2o)[k] + (*A2)[k] / (P->D2[k] + P->omega);
    }
    rms = sqrt(r1 + r2);
    niter = iter;
    if (rms < conv) break;
    if (do_diis) diis_step_x(s, *A1, *A1o, *A2, *A2o, 8, 0, min_vecs);
  }
  diis_free(s);
  if (out) {
    out[0] = orc_pert_pseudoresponse(P, hand);
    out[1] = rms;
    out[2] = niter;
  }
  return rms < conv ? 0 : 1;
}

void orc_pert_get(void *q, double *X1, double *X2, double *Y1, double *Y2) {
  orc_pert *P = (orc_pert *)q;
  const size_t ov = (size_t)P->s->no * P->s->nv;
  if (X1) memcpy(X1, P->X1, ov * sizeof(double));
  if (X2) memcpy(X2, P->X2, ov * ov * sizeof(double));
  if (Y1) memcpy(Y1, P->Y1, ov * sizeof(double));
  if (Y2) memcpy(Y2, P->Y2, ov * ov * sizeof(double));
}

void orc_pert_set(void *q, const double *X1, const double *X2, const double *Y1, const double *Y2) {
  orc_pert *P = (orc_pert *)q;
  const size_t ov = (size_t)P->s->no * P->s->nv;
  if (X1) memcpy(P->X1, X1, ov * sizeof(double));
  if (X2) memcpy(P->X2, X2, ov * ov * sizeof(double));
  if (Y1) memcpy(P->Y1, Y1, ov * sizeof(double));
  if (Y2) memcpy(P->Y2, Y2, ov * ov * sizeof(double));
}

/* CCResp::linresp, ccresp.cc:27-124: <<A;B>> from the perturbation of qx
 * (A, its pertbar blocks) and the amplitudes of qy (X of B and the Y of B in
 * its current state: the guess when only the right-hand equations were
 * solved), with the Lambda amplitudes */
double orc_linresp(void *qx, void *qy) {
  orc_pert *P = (orc_pert *)qx, *Q = (orc_pert *)qy;
  orc_state *s = P->s;
  const int no = s->no, nv = s->nv;
  const double *l1 = s->l1, *l2 = s->l2;
  const double *X1y = Q->X1, *X2y = Q->X2, *Y1y = Q->Y1, *Y2y = Q->Y2;
  double polar1 = 0.0, polar2 = 0.0;
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      polar1 += AVO(a, i) * Y1y[O2(i, a)];
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++) polar1 += 0.5 * (AVVOO(a, b, i, j) + AVVOO(b, a, j, i)) * Y2y[T4(i, j, a, b)];
    }
  for (int i = 0; i < no; i++)
    for (int a = 0; a < nv; a++) {
      const double la = l1[O2(i, a)];
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++)
          polar2 += la * PA(j, b + no) * (2.0 * X2y[T4(i, j, a, b)] - X2y[T4(i, j, b, a)]);
      for (int c = 0; c < nv; c++) polar2 += la * AVV(a, c) * X1y[O2(i, c)];
      for (int k = 0; k < no; k++) polar2 -= la * AOO(k, i) * X1y[O2(k, a)];
      for (int j = 0; j < no; j++)
        for (int b = 0; b < nv; b++) {
          const double l = l2[T4(i, j, a, b)];
          for (int c = 0; c < nv; c++) polar2 += l2[T4(i, j, b, c)] * X1y[O2(i, a)] * AVVVO(b, c, a, j);
          for (int k = 0; k < no; k++) {
            polar2 -= 0.5 * l * X1y[O2(k, a)] * AOVOO(k, b, i, j);
            polar2 -= 0.5 * l * X1y[O2(k, b)] * AOVOO(k, a, j, i);
            polar2 -= 0.5 * l * X2y[T4(k, j, a, b)] * AOO(k, i);
            polar2 -= 0.5 * l * X2y[T4(k, i, b, a)] * AOO(k, j);
          }
          for (int c = 0; c < nv; c++) {
            polar2 += 0.5 * l * X2y[T4(i, j, a, c)] * AVV(b, c);
            polar2 += 0.5 * l * X2y[T4(j, i, b, c)] * AVV(a, c);
          }
        }
      polar2 += 2.0 * PA(i, a + no) * X1y[O2(i, a)];
    }
  return -1.0 * (polar1 + polar2);
}
#undef Y1_
#undef Y2_
#undef X1_
#undef X2_
#undef l1_
#undef l2_
#undef T2_
#undef GVV
#undef GOO
#undef LV
#undef PA
#undef AOV
#undef AOO
#undef AVV
#undef AVO
#undef AOVOO
#undef AVVVO
#undef AVVOO
#undef HB
#undef Q4
