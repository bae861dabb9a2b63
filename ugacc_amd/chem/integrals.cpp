// Gaussian integrals over s, p, d shells (McMurchie-Davidson), host side.
//
// This is the integral part of the chemistry front end that replaces what
// the reference takes from Psi4 (libmints one-electron integrals and the
// libtrans two-electron integrals, hamiltonian.cc:44-94).  It computes, in a
// real spherical-harmonic basis (pure d, 5 components), the overlap S,
// kinetic T and nuclear attraction V matrices and the full AO two-electron
// array (pq|rs) (chemist notation), exploiting the 8-fold permutational
// symmetry and Schwarz screening.
//
// Method: Hermite expansion of Gaussian overlap distributions,
//   E^{ij}_0 = exp(-mu X_AB^2),
//   E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1},
// Hermite Coulomb integrals R^n_{tuv} from the Boys function
//   R^n_{000} = (-2a)^n F_n(a R^2),
//   R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v} (u, v alike),
// and (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) sum E^{ab}_{tuv} (-1)^{tau+nu+phi}
// E^{cd}_{tau nu phi} R_{t+tau, u+nu, v+phi}.
// The Boys function uses a tabulated grid with an 8-term Taylor expansion
// (T < 40) and the asymptotic form with upward recursion beyond.
//
// C ABI (called from ugacc_amd/chem/integrals.py):
//   shells are given as arrays: center xyz (3*nsh), l (nsh), nprim (nsh),
//   offsets into exps/coefs (nsh); coefficients refer to normalised
//   primitives and the contraction is renormalised here.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int LMAX = 2;
constexpr int LTOT = 4 * LMAX;  // highest Hermite order in an ERI
constexpr double PI = 3.14159265358979323846;

// ------------------------------------------------------------ Boys function
constexpr int BOYS_NMAX = LTOT + 8;
constexpr double BOYS_DT = 0.025;
constexpr double BOYS_TMAX = 40.0;
constexpr int BOYS_NGRID = (int)(BOYS_TMAX / BOYS_DT) + 2;
std::vector<double> g_boys;  // [grid][n], n = 0..BOYS_NMAX + 8

double boys_series(int n, double T) {
  // F_n(T) = e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))
  double term = 1.0 / (2 * n + 1), sum = term;
  for (int k = 1; k < 2000; k++) {
    term *= 2.0 * T / (2 * n + 2 * k + 1);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return std::exp(-T) * sum;
}

void boys_init() {
  if (!g_boys.empty()) return;
  const int NN = BOYS_NMAX + 9;
  std::vector<double> tab((size_t)BOYS_NGRID * NN);
  for (int g = 0; g < BOYS_NGRID; g++) {
    double T = g * BOYS_DT;
    // series at the top order, downward recursion (stable) for the rest
    tab[(size_t)g * NN + NN - 1] = boys_series(NN - 1, T);
    double e = std::exp(-T);
    for (int n = NN - 1; n > 0; n--) tab[(size_t)g * NN + n - 1] = (2 * T * tab[(size_t)g * NN + n] + e) / (2 * n - 1);
  }
  g_boys.swap(tab);
}

// F_0..F_nmax at T
inline void boys(int nmax, double T, double *F) {
  if (T < BOYS_TMAX) {
    const int NN = BOYS_NMAX + 9;
    int g = (int)(T / BOYS_DT + 0.5);
    double d = g * BOYS_DT - T;  // expansion about the nearest grid point
    const double *row = &g_boys[(size_t)g * NN];
    // F_n(T) = sum_k F_{n+k}(Tg) d^k / k!  (d/dT F_n = -F_{n+1})
    double e = std::exp(-T);
    double s = 0.0, dk = 1.0;
    for (int k = 0; k < 9; k++) {
      s += row[nmax + k] * dk;
      dk *= d / (k + 1);
    }
    F[nmax] = s;
    for (int n = nmax; n > 0; n--) F[n - 1] = (2 * T * F[n] + e) / (2 * n - 1);
  } else {
    double e = std::exp(-T);
    F[0] = 0.5 * std::sqrt(PI / T);
    for (int n = 0; n < nmax; n++) F[n + 1] = ((2 * n + 1) * F[n] - e) / (2 * T);
  }
}

// ---------------------------------------------------------------- shells
struct Shell {
  double A[3];
  int l;
  std::vector<double> a, c;  // exponents, contraction coefs incl. normalisation of x^l primitive
  int ncart, nsph, off;      // basis offset in the spherical basis
};

inline int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }
inline int nsph_of(int l) { return 2 * l + 1; }

// cartesian components of angular momentum l, ordered xx, xy, xz, yy, yz, zz
void cart_list(int l, int (*lmn)[3]) {
  int k = 0;
  for (int i = l; i >= 0; i--)
    for (int j = l - i; j >= 0; j--) {
      lmn[k][0] = i; lmn[k][1] = j; lmn[k][2] = l - i - j;
      k++;
    }
}

double dfact(int n) {  // (n)!! with (-1)!! = 1
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// Cartesian -> real solid harmonic coefficients: sph[s] = sum_c T[s][c] cart[c],
// cartesians all normalised like x^l (so x^l itself has unit norm).
void cart2sph(int l, double (*T)[6]) {
  for (int s = 0; s < 5; s++)
    for (int c = 0; c < 6; c++) T[s][c] = 0.0;
  if (l == 0) { T[0][0] = 1.0; return; }
  if (l == 1) { T[0][0] = 1.0; T[1][1] = 1.0; T[2][2] = 1.0; return; }
  // order: xx xy xz yy yz zz
  const double r3 = std::sqrt(3.0);
  T[0][5] = 1.0; T[0][0] = -0.5; T[0][3] = -0.5;   // d0  = zz - (xx + yy)/2
  T[1][2] = r3;                                   // d+1 = sqrt3 xz
  T[2][4] = r3;                                   // d-1 = sqrt3 yz
  T[3][0] = 0.5 * r3; T[3][3] = -0.5 * r3;        // d+2 = sqrt3/2 (xx - yy)
  T[4][1] = r3;                                   // d-2 = sqrt3 xy
}

// ------------------------------------------------------ Hermite E tables
// E[i][j][t] for i <= la, j <= lb, t <= i+j along one axis
struct E1 {
  double e[LMAX + 1][LMAX + 1 + 2][2 * LMAX + 3];
};

void hermite_1d(int la, int lb, double a, double b, double XA, double XB, E1 &E) {
  // XA = A - B along the axis
  double p = a + b, mu = a * b / p;
  double XPA = -b / p * XA, XPB = a / p * XA;  // P - A, P - B
  std::memset(&E, 0, sizeof(E));
  E.e[0][0][0] = std::exp(-mu * XA * XA);
  double ip2 = 0.5 / p;
  for (int i = 0; i <= la; i++)
    for (int j = 0; j <= lb; j++) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; t++) {
        double v;
        if (i > 0) {
          const double *src = E.e[i - 1][j];
          v = (t > 0 ? ip2 * src[t - 1] : 0.0) + XPA * src[t] + (t + 1 <= i - 1 + j ? (t + 1) * src[t + 1] : 0.0);
        } else {
          const double *src = E.e[i][j - 1];
          v = (t > 0 ? ip2 * src[t - 1] : 0.0) + XPB * src[t] + (t + 1 <= i + j - 1 ? (t + 1) * src[t + 1] : 0.0);
        }
        E.e[i][j][t] = v;
      }
    }
  (void)XB;
}

// --------------------------------------------------- shell-pair data
struct PrimPair {
  double p, P[3], coef;  // coef = ca * cb
  // Hermite coefficients per cartesian component pair: [ca*ncb + cb][tuv]
  std::vector<double> Etuv;
};
struct ShellPair {
  int sa, sb, L, ntuv;
  std::vector<PrimPair> prims;
  std::vector<int> tuv;  // packed (t,u,v) triples, t+u+v <= L
};

int tuv_count(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

void make_pair(const std::vector<Shell> &sh, int sa, int sb, ShellPair &sp) {
  const Shell &A = sh[sa], &B = sh[sb];
  sp.sa = sa; sp.sb = sb; sp.L = A.l + B.l;
  sp.ntuv = tuv_count(sp.L);
  sp.tuv.clear();
  for (int n = 0; n <= sp.L; n++)
    for (int t = n; t >= 0; t--)
      for (int u = n - t; u >= 0; u--) {
        sp.tuv.push_back(t); sp.tuv.push_back(u); sp.tuv.push_back(n - t - u);
      }
  int la[6][3], lb[6][3];
  cart_list(A.l, la);
  cart_list(B.l, lb);
  const int nca = ncart_of(A.l), ncb = ncart_of(B.l);
  sp.prims.clear();
  for (size_t i = 0; i < A.a.size(); i++)
    for (size_t j = 0; j < B.a.size(); j++) {
      PrimPair pp;
      double a = A.a[i], b = B.a[j];
      pp.p = a + b;
      for (int x = 0; x < 3; x++) pp.P[x] = (a * A.A[x] + b * B.A[x]) / pp.p;
      pp.coef = A.c[i] * B.c[j];
      E1 Ex, Ey, Ez;
      hermite_1d(A.l, B.l, a, b, A.A[0] - B.A[0], 0, Ex);
      hermite_1d(A.l, B.l, a, b, A.A[1] - B.A[1], 0, Ey);
      hermite_1d(A.l, B.l, a, b, A.A[2] - B.A[2], 0, Ez);
      pp.Etuv.assign((size_t)nca * ncb * sp.ntuv, 0.0);
      for (int ca = 0; ca < nca; ca++)
        for (int cb = 0; cb < ncb; cb++)
          for (int k = 0; k < sp.ntuv; k++) {
            int t = sp.tuv[3 * k], u = sp.tuv[3 * k + 1], v = sp.tuv[3 * k + 2];
            double ex = (t <= la[ca][0] + lb[cb][0]) ? Ex.e[la[ca][0]][lb[cb][0]][t] : 0.0;
            double ey = (u <= la[ca][1] + lb[cb][1]) ? Ey.e[la[ca][1]][lb[cb][1]][u] : 0.0;
            double ez = (v <= la[ca][2] + lb[cb][2]) ? Ez.e[la[ca][2]][lb[cb][2]][v] : 0.0;
            pp.Etuv[((size_t)ca * ncb + cb) * sp.ntuv + k] = ex * ey * ez;
          }
      sp.prims.push_back(std::move(pp));
    }
}

// R_{tuv} for t+u+v <= L at exponent alpha and vector X (Hermite Coulomb)
struct RTab {
  double r[LTOT + 1][LTOT + 1][LTOT + 1];
};

void hermite_R(int L, double alpha, const double *X, RTab &R) {
  double F[LTOT + 2];
  double R2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  boys(L, alpha * R2, F);
  // Rn[n][t][u][v] built for decreasing n
  static thread_local double Rn[LTOT + 1][LTOT + 1][LTOT + 1][LTOT + 1];
  double m2a = -2.0 * alpha, f = 1.0;
  for (int n = 0; n <= L; n++) {
    Rn[n][0][0][0] = f * F[n];
    f *= m2a;
  }
  for (int n = L - 1; n >= 0; n--) {
    int top = L - n;  // t+u+v <= top at this n
    for (int t = 0; t <= top; t++)
      for (int u = 0; u <= top - t; u++)
        for (int v = 0; v <= top - t - u; v++) {
          if (t + u + v == 0) continue;
          double val;
          if (t > 0) val = (t > 1 ? (t - 1) * Rn[n + 1][t - 2][u][v] : 0.0) + X[0] * Rn[n + 1][t - 1][u][v];
          else if (u > 0) val = (u > 1 ? (u - 1) * Rn[n + 1][t][u - 2][v] : 0.0) + X[1] * Rn[n + 1][t][u - 1][v];
          else val = (v > 1 ? (v - 1) * Rn[n + 1][t][u][v - 2] : 0.0) + X[2] * Rn[n + 1][t][u][v - 1];
          Rn[n][t][u][v] = val;
        }
  }
  for (int t = 0; t <= L; t++)
    for (int u = 0; u <= L - t; u++)
      for (int v = 0; v <= L - t - u; v++) R.r[t][u][v] = Rn[0][t][u][v];
}

// Cartesian quartet (nca*ncb*ncc*ncd, row-major) for a shell quartet
void eri_quartet_cart(const ShellPair &ab, const ShellPair &cd, int nca, int ncb, int ncc, int ncd,
                      std::vector<double> &out, std::vector<double> &G) {
  const int nab = nca * ncb, ncdn = ncc * ncd;
  const int Lt = ab.L + cd.L;
  out.assign((size_t)nab * ncdn, 0.0);
  G.resize((size_t)ncdn * ab.ntuv);
  RTab R;
  for (const PrimPair &P1 : ab.prims)
    for (const PrimPair &P2 : cd.prims) {
      double p = P1.p, q = P2.p;
      double alpha = p * q / (p + q);
      double X[3] = {P1.P[0] - P2.P[0], P1.P[1] - P2.P[1], P1.P[2] - P2.P[2]};
      hermite_R(Lt, alpha, X, R);
      double pref = 2.0 * std::pow(PI, 2.5) / (p * q * std::sqrt(p + q)) * P1.coef * P2.coef;
      // G[cd][tuv_ab] = sum_{tau nu phi} (-1)^{tau+nu+phi} E^cd R_{t+tau,...}
      for (int c = 0; c < ncdn; c++) {
        const double *Ecd = &P2.Etuv[(size_t)c * cd.ntuv];
        for (int k = 0; k < ab.ntuv; k++) {
          int t = ab.tuv[3 * k], u = ab.tuv[3 * k + 1], v = ab.tuv[3 * k + 2];
          double s = 0.0;
          for (int m = 0; m < cd.ntuv; m++) {
            int tt = cd.tuv[3 * m], uu = cd.tuv[3 * m + 1], vv = cd.tuv[3 * m + 2];
            double e = Ecd[m];
            if (e == 0.0) continue;
            double r = R.r[t + tt][u + uu][v + vv];
            s += ((tt + uu + vv) & 1) ? -e * r : e * r;
          }
          G[(size_t)c * ab.ntuv + k] = s;
        }
      }
      for (int a = 0; a < nab; a++) {
        const double *Eab = &P1.Etuv[(size_t)a * ab.ntuv];
        for (int c = 0; c < ncdn; c++) {
          const double *g = &G[(size_t)c * ab.ntuv];
          double s = 0.0;
          for (int k = 0; k < ab.ntuv; k++) s += Eab[k] * g[k];
          out[(size_t)a * ncdn + c] += pref * s;
        }
      }
    }
}

struct Basis {
  std::vector<Shell> sh;
  int nbf = 0;
};

Basis make_basis(int nsh, const double *xyz, const int *l, const int *nprim, const int *off, const double *exps,
                 const double *coefs) {
  Basis B;
  int o = 0;
  for (int s = 0; s < nsh; s++) {
    Shell S;
    for (int x = 0; x < 3; x++) S.A[x] = xyz[3 * s + x];
    S.l = l[s];
    S.ncart = ncart_of(S.l);
    S.nsph = nsph_of(S.l);
    S.off = o;
    o += S.nsph;
    double dfl = dfact(2 * S.l - 1);
    for (int k = 0; k < nprim[s]; k++) {
      double a = exps[off[s] + k];
      // normalisation of x^l exp(-a r^2)
      double N = std::pow(2 * a / PI, 0.75) * std::pow(4 * a, 0.5 * S.l) / std::sqrt(dfl);
      S.a.push_back(a);
      S.c.push_back(coefs[off[s] + k] * N);
    }
    // renormalise the contraction of x^l
    double ss = 0.0;
    for (size_t i = 0; i < S.a.size(); i++)
      for (size_t j = 0; j < S.a.size(); j++) {
        double p = S.a[i] + S.a[j];
        ss += S.c[i] * S.c[j] * std::pow(PI / p, 1.5) * dfl / std::pow(2 * p, S.l);
      }
    for (double &c : S.c) c /= std::sqrt(ss);
    B.sh.push_back(S);
  }
  B.nbf = o;
  return B;
}

// one-electron cartesian block for shells A, B: S, T, V
void one_cart(const Shell &A, const Shell &B, int natom, const double *Z, const double *C, double *S, double *T,
              double *V) {
  int la[6][3], lb[6][3];
  cart_list(A.l, la);
  cart_list(B.l, lb);
  const int nca = ncart_of(A.l), ncb = ncart_of(B.l);
  for (int k = 0; k < nca * ncb; k++) S[k] = T[k] = V[k] = 0.0;
  for (size_t i = 0; i < A.a.size(); i++)
    for (size_t j = 0; j < B.a.size(); j++) {
      double a = A.a[i], b = B.a[j], p = a + b;
      double cc = A.c[i] * B.c[j];
      E1 E[3];
      // kinetic needs j up to lb + 2
      int lbk = B.l + 2;
      double P[3];
      for (int x = 0; x < 3; x++) {
        P[x] = (a * A.A[x] + b * B.A[x]) / p;
        // reuse hermite_1d with lb extended: build into a larger table
      }
      // 1D overlaps s[x][i][j] for j <= lb+2
      double s1[3][LMAX + 1][LMAX + 3];
      for (int x = 0; x < 3; x++) {
        double XA = A.A[x] - B.A[x];
        double mu = a * b / p, XPA = P[x] - A.A[x], XPB = P[x] - B.A[x];
        double e[LMAX + 2][LMAX + 4][2 * LMAX + 6];
        std::memset(e, 0, sizeof(e));
        e[0][0][0] = std::exp(-mu * XA * XA);
        double ip2 = 0.5 / p;
        for (int ii = 0; ii <= A.l; ii++)
          for (int jj = 0; jj <= lbk; jj++) {
            if (ii == 0 && jj == 0) continue;
            for (int t = 0; t <= ii + jj; t++) {
              double v;
              if (ii > 0) {
                double *src = e[ii - 1][jj];
                v = (t > 0 ? ip2 * src[t - 1] : 0.0) + XPA * src[t] + (t + 1 <= ii - 1 + jj ? (t + 1) * src[t + 1] : 0.0);
              } else {
                double *src = e[ii][jj - 1];
                v = (t > 0 ? ip2 * src[t - 1] : 0.0) + XPB * src[t] + (t + 1 <= jj - 1 ? (t + 1) * src[t + 1] : 0.0);
              }
              e[ii][jj][t] = v;
            }
          }
        for (int ii = 0; ii <= A.l; ii++)
          for (int jj = 0; jj <= lbk; jj++) s1[x][ii][jj] = e[ii][jj][0] * std::sqrt(PI / p);
        for (int ii = 0; ii <= A.l; ii++)
          for (int jj = 0; jj <= B.l; jj++)
            for (int t = 0; t <= ii + jj; t++) E[x].e[ii][jj][t] = e[ii][jj][t];
      }
      for (int ca = 0; ca < nca; ca++)
        for (int cb = 0; cb < ncb; cb++) {
          double sx = s1[0][la[ca][0]][lb[cb][0]], sy = s1[1][la[ca][1]][lb[cb][1]], sz = s1[2][la[ca][2]][lb[cb][2]];
          double k3[3];
          for (int x = 0; x < 3; x++) {
            int ii = la[ca][x], jj = lb[cb][x];
            double t = -2.0 * b * b * s1[x][ii][jj + 2] + b * (2 * jj + 1) * s1[x][ii][jj];
            if (jj >= 2) t -= 0.5 * jj * (jj - 1) * s1[x][ii][jj - 2];
            k3[x] = t;
          }
          S[ca * ncb + cb] += cc * sx * sy * sz;
          T[ca * ncb + cb] += cc * (k3[0] * sy * sz + sx * k3[1] * sz + sx * sy * k3[2]);
        }
      // nuclear attraction
      const int L = A.l + B.l;
      for (int n = 0; n < natom; n++) {
        double X[3] = {P[0] - C[3 * n], P[1] - C[3 * n + 1], P[2] - C[3 * n + 2]};
        RTab R;
        hermite_R(L, p, X, R);
        double pref = -Z[n] * 2.0 * PI / p * cc;
        for (int ca = 0; ca < nca; ca++)
          for (int cb = 0; cb < ncb; cb++) {
            double s = 0.0;
            int ix = la[ca][0], jx = lb[cb][0], iy = la[ca][1], jy = lb[cb][1], iz = la[ca][2], jz = lb[cb][2];
            for (int t = 0; t <= ix + jx; t++)
              for (int u = 0; u <= iy + jy; u++)
                for (int v = 0; v <= iz + jz; v++)
                  s += E[0].e[ix][jx][t] * E[1].e[iy][jy][u] * E[2].e[iz][jz][v] * R.r[t][u][v];
            V[ca * ncb + cb] += pref * s;
          }
      }
    }
}

// 1-D overlaps s[i][j] = <x_A^i e^(-a x_A^2) | x_B^j e^(-b x_B^2)> for i <= la,
// j <= jmax (McMurchie-Davidson E^{ij}_0 times sqrt(pi / p))
void overlap_1d(double a, double b, double Ax, double Bx, int la, int jmax, double (*s)[LMAX + 4]) {
  const double p = a + b, mu = a * b / p, Px = (a * Ax + b * Bx) / p;
  const double XPA = Px - Ax, XPB = Px - Bx, ip2 = 0.5 / p;
  double e[LMAX + 2][LMAX + 4][2 * LMAX + 6];
  std::memset(e, 0, sizeof(e));
  e[0][0][0] = std::exp(-mu * (Ax - Bx) * (Ax - Bx));
  for (int i = 0; i <= la; i++)
    for (int j = 0; j <= jmax; j++) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; t++) {
        const double *src = i > 0 ? e[i - 1][j] : e[i][j - 1];
        const int top = i > 0 ? i - 1 + j : j - 1;
        e[i][j][t] = (t > 0 ? ip2 * src[t - 1] : 0.0) + (i > 0 ? XPA : XPB) * src[t] +
                     (t + 1 <= top ? (t + 1) * src[t + 1] : 0.0);
      }
    }
  const double f = std::sqrt(PI / p);
  for (int i = 0; i <= la; i++)
    for (int j = 0; j <= jmax; j++) s[i][j] = e[i][j][0] * f;
}

// Property integrals of a cartesian shell pair about the point O:
// D[x] = <A| (r - O)_x |B> and M[x] = <A| ((r - O) x grad)_x |B> (real,
// antisymmetric: M[0] = <A| y_O d/dz - z_O d/dy |B>, ...).  Per axis, with
// B's power j about B: (x - O_x) B_j = B_{j+1} + (B_x - O_x) B_j and
// d/dx B_j = j B_{j-1} - 2 b B_{j+1}.
void prop_cart(const Shell &A, const Shell &B, const double *O, double *D, double *M) {
  int la[6][3], lb[6][3];
  cart_list(A.l, la);
  cart_list(B.l, lb);
  const int nca = ncart_of(A.l), ncb = ncart_of(B.l);
  for (int x = 0; x < 3; x++)
    for (int k = 0; k < nca * ncb; k++) D[x * 36 + k] = M[x * 36 + k] = 0.0;
  for (size_t i = 0; i < A.a.size(); i++)
    for (size_t j = 0; j < B.a.size(); j++) {
      const double a = A.a[i], b = B.a[j], cc = A.c[i] * B.c[j];
      double s[3][LMAX + 2][LMAX + 4];
      for (int x = 0; x < 3; x++) overlap_1d(a, b, A.A[x], B.A[x], A.l, B.l + 1, s[x]);
      for (int ca = 0; ca < nca; ca++)
        for (int cb = 0; cb < ncb; cb++) {
          double ov[3], mo[3], de[3];
          for (int x = 0; x < 3; x++) {
            const int ii = la[ca][x], jj = lb[cb][x];
            ov[x] = s[x][ii][jj];
            mo[x] = s[x][ii][jj + 1] + (B.A[x] - O[x]) * s[x][ii][jj];
            de[x] = (jj > 0 ? jj * s[x][ii][jj - 1] : 0.0) - 2.0 * b * s[x][ii][jj + 1];
          }
          const int k = ca * ncb + cb;
          D[0 * 36 + k] += cc * mo[0] * ov[1] * ov[2];
          D[1 * 36 + k] += cc * ov[0] * mo[1] * ov[2];
          D[2 * 36 + k] += cc * ov[0] * ov[1] * mo[2];
          M[0 * 36 + k] += cc * ov[0] * (mo[1] * de[2] - de[1] * mo[2]);
          M[1 * 36 + k] += cc * ov[1] * (mo[2] * de[0] - de[2] * mo[0]);
          M[2 * 36 + k] += cc * ov[2] * (mo[0] * de[1] - de[0] * mo[1]);
        }
    }
}

// cartesian block (na_c x nb_c) -> spherical (na_s x nb_s)
void to_sph2(int la, int lb, const double *cart, double *sph) {
  double Ta[5][6], Tb[5][6];
  cart2sph(la, Ta);
  cart2sph(lb, Tb);
  int nca = ncart_of(la), ncb = ncart_of(lb), nsa = nsph_of(la), nsb = nsph_of(lb);
  for (int i = 0; i < nsa; i++)
    for (int j = 0; j < nsb; j++) {
      double s = 0.0;
      for (int a = 0; a < nca; a++) {
        if (Ta[i][a] == 0.0) continue;
        for (int b = 0; b < ncb; b++) s += Ta[i][a] * Tb[j][b] * cart[a * ncb + b];
      }
      sph[i * nsb + j] = s;
    }
}

}  // namespace

extern "C" {

int chem_nbf(int nsh, const int *l) {
  int n = 0;
  for (int s = 0; s < nsh; s++) n += nsph_of(l[s]);
  return n;
}

// S, T, V: nbf x nbf each (row-major)
int chem_one_electron(int nsh, const double *xyz, const int *l, const int *nprim, const int *off, const double *exps,
                      const double *coefs, int natom, const double *Z, const double *C, double *S, double *T,
                      double *V) {
  for (int s = 0; s < nsh; s++)
    if (l[s] < 0 || l[s] > LMAX) return -1;
  boys_init();
  Basis B = make_basis(nsh, xyz, l, nprim, off, exps, coefs);
  const int n = B.nbf;
#pragma omp parallel for schedule(dynamic)
  for (int a = 0; a < nsh; a++)
    for (int b = 0; b < nsh; b++) {
      const Shell &A = B.sh[a], &Bs = B.sh[b];
      double cs[36], ct[36], cv[36], ss[25], st[25], sv[25];
      one_cart(A, Bs, natom, Z, C, cs, ct, cv);
      to_sph2(A.l, Bs.l, cs, ss);
      to_sph2(A.l, Bs.l, ct, st);
      to_sph2(A.l, Bs.l, cv, sv);
      for (int i = 0; i < A.nsph; i++)
        for (int j = 0; j < Bs.nsph; j++) {
          size_t k = (size_t)(A.off + i) * n + Bs.off + j;
          S[k] = ss[i * Bs.nsph + j];
          T[k] = st[i * Bs.nsph + j];
          V[k] = sv[i * Bs.nsph + j];
        }
    }
  return 0;
}

// Dipole and angular-momentum integrals about origin[3]: D[3][nbf][nbf] =
// <p| (r - O) |q>, M[3][nbf][nbf] = <p| (r - O) x grad |q> (real parts of
// the operators, without charges or factors of -i)
int chem_properties(int nsh, const double *xyz, const int *l, const int *nprim, const int *off, const double *exps,
                    const double *coefs, const double *origin, double *D, double *M) {
  for (int s = 0; s < nsh; s++)
    if (l[s] < 0 || l[s] > LMAX) return -1;
  Basis B = make_basis(nsh, xyz, l, nprim, off, exps, coefs);
  const long n = B.nbf;
#pragma omp parallel for schedule(dynamic)
  for (int a = 0; a < nsh; a++)
    for (int b = 0; b < nsh; b++) {
      const Shell &A = B.sh[a], &Bs = B.sh[b];
      double cd[3 * 36], cm[3 * 36], sd[25], sm[25];
      prop_cart(A, Bs, origin, cd, cm);
      for (int x = 0; x < 3; x++) {
        to_sph2(A.l, Bs.l, cd + 36 * x, sd);
        to_sph2(A.l, Bs.l, cm + 36 * x, sm);
        for (int i = 0; i < A.nsph; i++)
          for (int j = 0; j < Bs.nsph; j++) {
            const size_t k = (size_t)x * n * n + (size_t)(A.off + i) * n + Bs.off + j;
            D[k] = sd[i * Bs.nsph + j];
            M[k] = sm[i * Bs.nsph + j];
          }
      }
    }
  return 0;
}

// Full AO ERI array (pq|rs), nbf^4 doubles, row-major.  Quartets whose
// Schwarz bound is below thresh are set to zero.
int chem_eri(int nsh, const double *xyz, const int *l, const int *nprim, const int *off, const double *exps,
             const double *coefs, double thresh, double *eri) {
  for (int s = 0; s < nsh; s++)
    if (l[s] < 0 || l[s] > LMAX) return -1;
  boys_init();
  Basis B = make_basis(nsh, xyz, l, nprim, off, exps, coefs);
  const long n = B.nbf;
  // shell pairs a >= b
  std::vector<std::pair<int, int>> pairs;
  for (int a = 0; a < nsh; a++)
    for (int b = 0; b <= a; b++) pairs.push_back({a, b});
  const long np = (long)pairs.size();
  std::vector<ShellPair> sp(np);
#pragma omp parallel for schedule(dynamic)
  for (long k = 0; k < np; k++) make_pair(B.sh, pairs[k].first, pairs[k].second, sp[k]);
  // Schwarz factors: sqrt(max |(ab|ab)|)
  std::vector<double> qab(np);
#pragma omp parallel
  {
    std::vector<double> out, G;
#pragma omp for schedule(dynamic)
    for (long k = 0; k < np; k++) {
      const Shell &A = B.sh[pairs[k].first], &Bs = B.sh[pairs[k].second];
      eri_quartet_cart(sp[k], sp[k], A.ncart, Bs.ncart, A.ncart, Bs.ncart, out, G);
      double mx = 0.0;
      int nab = A.ncart * Bs.ncart;
      for (int i = 0; i < nab; i++) mx = std::fmax(mx, std::fabs(out[(size_t)i * nab + i]));
      qab[k] = std::sqrt(mx) * 3.0;  // margin for the cartesian -> spherical mixing
    }
  }
#pragma omp parallel
  {
    std::vector<double> out, G, tmp1, tmp2;
#pragma omp for schedule(dynamic, 1)
    for (long k1 = 0; k1 < np; k1++) {
      for (long k2 = 0; k2 <= k1; k2++) {
        if (qab[k1] * qab[k2] < thresh) continue;
        const int a = pairs[k1].first, b = pairs[k1].second, c = pairs[k2].first, d = pairs[k2].second;
        const Shell &A = B.sh[a], &Bs = B.sh[b], &Cs = B.sh[c], &D = B.sh[d];
        eri_quartet_cart(sp[k1], sp[k2], A.ncart, Bs.ncart, Cs.ncart, D.ncart, out, G);
        // cartesian -> spherical on each index
        double Ta[5][6], Tb[5][6], Tc[5][6], Td[5][6];
        cart2sph(A.l, Ta); cart2sph(Bs.l, Tb); cart2sph(Cs.l, Tc); cart2sph(D.l, Td);
        const int ca = A.ncart, cb = Bs.ncart, cc = Cs.ncart, cd = D.ncart;
        const int sa = A.nsph, sb = Bs.nsph, sc = Cs.nsph, sd = D.nsph;
        // transform d
        tmp1.assign((size_t)ca * cb * cc * sd, 0.0);
        for (int x = 0; x < ca * cb * cc; x++)
          for (int j = 0; j < sd; j++) {
            double s = 0.0;
            for (int y = 0; y < cd; y++) s += Td[j][y] * out[(size_t)x * cd + y];
            tmp1[(size_t)x * sd + j] = s;
          }
        // transform c
        tmp2.assign((size_t)ca * cb * sc * sd, 0.0);
        for (int x = 0; x < ca * cb; x++)
          for (int i = 0; i < sc; i++)
            for (int j = 0; j < sd; j++) {
              double s = 0.0;
              for (int y = 0; y < cc; y++) s += Tc[i][y] * tmp1[((size_t)x * cc + y) * sd + j];
              tmp2[((size_t)x * sc + i) * sd + j] = s;
            }
        // transform b
        tmp1.assign((size_t)ca * sb * sc * sd, 0.0);
        for (int x = 0; x < ca; x++)
          for (int i = 0; i < sb; i++)
            for (int r = 0; r < sc * sd; r++) {
              double s = 0.0;
              for (int y = 0; y < cb; y++) s += Tb[i][y] * tmp2[((size_t)x * cb + y) * sc * sd + r];
              tmp1[((size_t)x * sb + i) * sc * sd + r] = s;
            }
        // transform a and scatter with 8-fold symmetry
        for (int i = 0; i < sa; i++)
          for (int r = 0; r < sb * sc * sd; r++) {
            double s = 0.0;
            for (int y = 0; y < ca; y++) s += Ta[i][y] * tmp1[(size_t)y * sb * sc * sd + r];
            int jb = r / (sc * sd), kc = (r / sd) % sc, ld = r % sd;
            long p = A.off + i, q = Bs.off + jb, rr = Cs.off + kc, ss = D.off + ld;
            eri[((p * n + q) * n + rr) * n + ss] = s;
            eri[((q * n + p) * n + rr) * n + ss] = s;
            eri[((p * n + q) * n + ss) * n + rr] = s;
            eri[((q * n + p) * n + ss) * n + rr] = s;
            eri[((rr * n + ss) * n + p) * n + q] = s;
            eri[((ss * n + rr) * n + p) * n + q] = s;
            eri[((rr * n + ss) * n + q) * n + p] = s;
            eri[((ss * n + rr) * n + q) * n + p] = s;
          }
      }
    }
  }
  return 0;
}

// Coulomb and exchange: J_pq = sum_rs (pq|rs) D_rs, K_pq = sum_rs (pr|qs) D_rs
void chem_jk(int n, const double *eri, const double *D, double *J, double *K) {
  const long N = n;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; p++)
    for (long q = 0; q < N; q++) {
      double sj = 0.0, sk = 0.0;
      const double *e1 = eri + ((p * N + q) * N) * N;
      for (long r = 0; r < N * N; r++) sj += e1[r] * D[r];
      for (long r = 0; r < N; r++) {
        const double *e2 = eri + ((p * N + r) * N + q) * N;
        const double *d = D + r * N;
        for (long s = 0; s < N; s++) sk += e2[s] * d[s];
      }
      J[p * N + q] = sj;
      K[p * N + q] = sk;
    }
}

int chem_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
