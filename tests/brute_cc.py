"""Closed-shell CCSD in the determinant basis, for pins of the oracle.

The N-electron, S_z = 0 sector of Fock space is spanned by pairs of alpha
and beta occupation strings.  Spin-free one-electron excitation operators
E_pq = a+_pa a_qa + a+_pb a_qb are dense matrices there, and the Hamiltonian
is assembled from a one-electron matrix h and Dirac integrals <pq|rs>:

    H = sum_pq h_pq E_pq + 1/2 sum_pqrs <pq|rs> (E_pr E_qs - delta_qr E_ps).

The cluster operator of the closed-shell amplitudes (the reference's t1[i][a]
and t2[i][j][a][b]) is T = sum t1_ia E_ai + 1/2 sum t2_ijab E_ai E_bj, and
the CCSD equations are the projections of exp(-T) H exp(T)|0> on the
spin-adapted singles E_ai|0> and doubles E_ai E_bj|0>, which span the same
space as the reference's biorthogonal projectors, so the solutions agree.
Any one-electron operator, Hermitian or not, can be added to H: that is what
the pins of the response for general perturbations use.  Sizes stay tiny
(two occupied and up to four virtual orbitals: at most 225 determinants).
"""
import itertools

import numpy as np
from scipy.linalg import expm


def _strings(n, k):
    return [sum(1 << p for p in c) for c in itertools.combinations(range(n), k)]


def _one_spin_excitations(n, strings):
    """ops[p][q]: list of (row, col, sign) of a+_p a_q on the strings."""
    index = {s: k for k, s in enumerate(strings)}
    ops = [[[] for _ in range(n)] for _ in range(n)]
    for col, s in enumerate(strings):
        for q in range(n):
            if not s >> q & 1:
                continue
            s1 = s & ~(1 << q)
            sign_q = (-1) ** bin(s & ((1 << q) - 1)).count("1")
            for p in range(n):
                if s1 >> p & 1:
                    continue
                s2 = s1 | (1 << p)
                sign_p = (-1) ** bin(s1 & ((1 << p) - 1)).count("1")
                ops[p][q].append((index[s2], col, sign_p * sign_q))
    return ops


class Sector:
    """The determinants of ``no`` doubly occupied orbitals among ``n``."""

    def __init__(self, n, no):
        self.n, self.no = n, no
        st = _strings(n, no)
        ns = len(st)
        self.dim = ns * ns
        ref = st.index((1 << no) - 1)
        self.ref = np.zeros(self.dim)
        self.ref[ref * ns + ref] = 1.0
        one = _one_spin_excitations(n, st)
        E = np.zeros((n, n, self.dim, self.dim))
        for p in range(n):
            for q in range(n):
                for r, c, s in one[p][q]:
                    # index alpha * ns + beta; a beta pair a+ a moves past
                    # the alpha creators twice, so only its in-string sign
                    # remains
                    for b in range(ns):
                        E[p, q, r * ns + b, c * ns + b] += s
                        E[p, q, b * ns + r, b * ns + c] += s
        self.E = E

    def one_body(self, A):
        """sum_pq A_pq E_pq."""
        return np.tensordot(A, self.E, axes=([0, 1], [0, 1]))

    def hamiltonian(self, h, g):
        """h one-electron, g Dirac <pq|rs>."""
        n = self.n
        H = self.one_body(h)
        M = np.tensordot(g, self.E, axes=([1, 3], [0, 1]))  # M[p, r] = sum_qs g_pqrs E_qs
        for p in range(n):
            for r in range(n):
                H += 0.5 * self.E[p, r] @ M[p, r]
        H -= 0.5 * self.one_body(np.einsum("pqqs->ps", g))
        return H

    def cluster(self, t1, t2):
        no, n = self.no, self.n
        E = self.E
        T = np.zeros((self.dim, self.dim))
        for i in range(no):
            for a in range(no, n):
                T += t1[i, a - no] * E[a, i]
        for i, j in itertools.product(range(no), repeat=2):
            for a, b in itertools.product(range(no, n), repeat=2):
                x = t2[i, j, a - no, b - no]
                if x:
                    T += 0.5 * x * E[a, i] @ E[b, j]
        return T

    def project(self, v):
        """(<0|E_ia v>, <0|E_jb E_ia v>) as (o, v) and (o, o, v, v) arrays."""
        no, n = self.no, self.n
        nv = n - no
        r1 = np.zeros((no, nv))
        r2 = np.zeros((no, no, nv, nv))
        bra1 = {}
        for i in range(no):
            for a in range(no, n):
                bra1[i, a] = self.E[a, i] @ self.ref
                r1[i, a - no] = bra1[i, a] @ v
        for i, j in itertools.product(range(no), repeat=2):
            for a, b in itertools.product(range(no, n), repeat=2):
                r2[i, j, a - no, b - no] = (self.E[b, j] @ bra1[i, a]) @ v
        return r1, r2


def core_hamiltonian(fock, g, no):
    """h with f_pq = h_pq + sum_i (2 <pi|qi> - <pi|iq>)."""
    o = slice(0, no)
    return fock - 2.0 * np.einsum("piqi->pq", g[:, o, :, o]) + np.einsum("piiq->pq", g[:, o, o, :])


class BruteCC:
    """CCSD of H + V (V any one-electron operator) in a Sector."""

    def __init__(self, fock, g, no):
        self.no = no
        self.nv = fock.shape[0] - no
        self.S = Sector(fock.shape[0], no)
        self.H = self.S.hamiltonian(core_hamiltonian(fock, g, no), g)

    def _unpack(self, x):
        no, nv = self.no, self.nv
        t1 = x[:no * nv].reshape(no, nv)
        t2 = x[no * nv:].reshape(no, no, nv, nv)
        return t1, 0.5 * (t2 + t2.transpose(1, 0, 3, 2))

    def residual(self, t1, t2, V=None):
        T = self.S.cluster(t1, t2)
        H = self.H if V is None else self.H + V
        Hbar = expm(-T) @ H @ expm(T)
        return self.S.project(Hbar @ self.S.ref)

    def energy(self, t1, t2, V=None):
        """<0|exp(-T) (H + V) exp(T)|0>, the total energy."""
        T = self.S.cluster(t1, t2)
        H = self.H if V is None else self.H + V
        ref = self.S.ref
        return ref @ (expm(-T) @ (H @ (expm(T) @ ref)))

    def solve(self, t1, t2, V=None, conv=1e-13, maxiter=60):
        """Chord iterations from (t1, t2): the finite-difference Jacobian of
        the first call is kept for later solves of nearby Hamiltonians (a
        small field V), which then converge linearly and fast."""
        x = np.concatenate([t1.ravel(), t2.ravel()])

        def F(y):
            r1, r2 = self.residual(*self._unpack(y), V=V)
            return np.concatenate([r1.ravel(), r2.ravel()])

        J = getattr(self, "_jac", None)
        for _ in range(maxiter):
            f = F(x)
            if np.max(np.abs(f)) < conv:
                return self._unpack(x)
            if J is None:
                h = 1e-7
                J = np.empty((f.size, x.size))
                for k in range(x.size):
                    e = np.zeros(x.size)
                    e[k] = h
                    J[:, k] = (F(x + e) - F(x - e)) / (2 * h)
            x = x + np.linalg.lstsq(J, -f, rcond=1e-12)[0]
        raise RuntimeError("brute-force CCSD did not converge")

    def keep_jacobian(self, t1, t2):
        """Stores the Jacobian at (t1, t2) for the chord iterations."""
        x = np.concatenate([t1.ravel(), t2.ravel()])
        h = 1e-7
        cols = []
        for k in range(x.size):
            e = np.zeros(x.size)
            e[k] = h
            rp = self.residual(*self._unpack(x + e))
            rm = self.residual(*self._unpack(x - e))
            cols.append(np.concatenate([(rp[0] - rm[0]).ravel(), (rp[1] - rm[1]).ravel()]) / (2 * h))
        self._jac = np.array(cols).T

    def excitation(self, X1, X2):
        """X|0> projected: <mu| X |0> for X = sum X1 E_ai + 1/2 sum X2 E_ai E_bj."""
        return self.S.project(self.S.cluster(X1, X2) @ self.S.ref)
