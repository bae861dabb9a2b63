// Driver for the sanitized build of the front end's integral code
// (ugacc_amd/chem/integrals.cpp; tests/test_sanitizers.py): overlap, kinetic,
// nuclear attraction, two-electron, dipole and angular-momentum integrals and
// the J/K build on a small water-like molecule with s, p and d shells (two
// primitives each), built with -fsanitize=address,undefined.  Checks the
// symmetries each matrix has and prints "OK".
#include <cmath>
#include <cstdio>
#include <vector>

extern "C" {
int chem_nbf(int nsh, const int *l);
int chem_one_electron(int nsh, const double *xyz, const int *l, const int *nprim, const int *off, const double *exps,
                      const double *coefs, int natom, const double *Z, const double *C, double *S, double *T,
                      double *V);
int chem_properties(int nsh, const double *xyz, const int *l, const int *nprim, const int *off, const double *exps,
                    const double *coefs, const double *origin, double *D, double *M);
int chem_eri(int nsh, const double *xyz, const int *l, const int *nprim, const int *off, const double *exps,
             const double *coefs, double thresh, double *eri);
void chem_jk(int n, const double *eri, const double *D, double *J, double *K);
}

int main() {
  // O at the origin (s, s, p, d), two H (s, p)
  const double O[3] = {0.0, 0.0, -0.12}, H1[3] = {0.0, 1.43, 0.98}, H2[3] = {0.0, -1.43, 0.98};
  std::vector<double> xyz;
  std::vector<int> l, nprim, off;
  std::vector<double> exps, coefs;
  auto shell = [&](const double *c, int am, double e0, double e1) {
    xyz.insert(xyz.end(), c, c + 3);
    l.push_back(am);
    nprim.push_back(2);
    off.push_back((int)exps.size());
    exps.push_back(e0); exps.push_back(e1);
    coefs.push_back(0.6); coefs.push_back(0.45);
  };
  shell(O, 0, 11.7, 2.3); shell(O, 0, 0.9, 0.3); shell(O, 1, 5.0, 1.1); shell(O, 2, 1.2, 0.4);
  shell(H1, 0, 3.4, 0.6); shell(H1, 1, 0.9, 0.3); shell(H2, 0, 3.4, 0.6); shell(H2, 1, 0.9, 0.3);
  const int nsh = (int)l.size(), n = chem_nbf(nsh, l.data());
  const size_t n2 = (size_t)n * n;
  std::vector<double> S(n2), T(n2), V(n2), D(3 * n2), M(3 * n2), eri(n2 * n2), P(n2), J(n2), K(n2);
  const double Z[3] = {8.0, 1.0, 1.0}, C[9] = {O[0], O[1], O[2], H1[0], H1[1], H1[2], H2[0], H2[1], H2[2]};
  const double origin[3] = {0.1, -0.2, 0.3};
  int rc = chem_one_electron(nsh, xyz.data(), l.data(), nprim.data(), off.data(), exps.data(), coefs.data(), 3, Z, C,
                             S.data(), T.data(), V.data());
  rc |= chem_properties(nsh, xyz.data(), l.data(), nprim.data(), off.data(), exps.data(), coefs.data(), origin, D.data(),
                        M.data());
  rc |= chem_eri(nsh, xyz.data(), l.data(), nprim.data(), off.data(), exps.data(), coefs.data(), 1e-14, eri.data());
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++) P[(size_t)p * n + q] = (p == q ? 1.0 : 0.0) + 0.01 * ((p + q) % 3);
  chem_jk(n, eri.data(), P.data(), J.data(), K.data());
  double asym = 0.0;
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++) {
      const size_t pq = (size_t)p * n + q, qp = (size_t)q * n + p;
      asym = std::fmax(asym, std::fabs(S[pq] - S[qp]) + std::fabs(T[pq] - T[qp]) + std::fabs(V[pq] - V[qp]));
      for (int k = 0; k < 3; k++)
        asym = std::fmax(asym, std::fabs(D[k * n2 + pq] - D[k * n2 + qp]) + std::fabs(M[k * n2 + pq] + M[k * n2 + qp]));
      for (int r = 0; r < n; r++)
        for (int s = 0; s < n; s++)  // (pq|rs) = (rs|pq)
          asym = std::fmax(asym, std::fabs(eri[pq * n2 + (size_t)r * n + s] - eri[((size_t)r * n + s) * n2 + pq]));
    }
  if (rc != 0 || !(asym < 1e-10)) {
    std::printf("failed: rc %d asymmetry %.3e\n", rc, asym);
    return 1;
  }
  std::printf("OK nbf %d max asymmetry %.3e S00 %.12f J00 %.12f\n", n, asym, S[0], J[0]);
  return 0;
}
