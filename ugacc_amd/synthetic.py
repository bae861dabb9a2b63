"""Seeded synthetic closed-shell Hamiltonians for parity tests and benches.

A Hamiltonian here is what the reference's Hamiltonian class holds
(hamiltonian.cc:44-102): an active-space Fock matrix ``fock`` (nact x nact,
occupied orbitals first) and the Dirac integrals ``<pq|rs>`` over the active
orbitals, which for real orbitals obey the 8-fold symmetry of the chemist
integrals ``(pr|qs)``.

The two-electron part is built as a Gram matrix over orbital pairs, so it is
positive semi-definite like a real electron-repulsion matrix; the Fock matrix
has occupied energies below the virtual ones, so the amplitude equations are
well conditioned.
"""
import numpy as np


def chemist_from_factors(B):
    """(pq|rs) = sum_K B[p,q,K] B[r,s,K] with B symmetric in (p,q)."""
    n = B.shape[0]
    Bs = 0.5 * (B + B.transpose(1, 0, 2))
    M = Bs.reshape(n * n, -1)
    return (M @ M.T).reshape(n, n, n, n)


def dirac_from_chemist(eri):
    """<pq|rs> = (pr|qs)."""
    return np.ascontiguousarray(eri.transpose(0, 2, 1, 3))


def random_hamiltonian(no, nv, seed=0, scale=0.08, rank=None, offdiag=0.0, fov=0.0, gap=0.5):
    """Returns (fock, ints) for ``no`` occupied and ``nv`` virtual orbitals.

    offdiag: size of random off-diagonal oo and vv Fock elements.
    fov:     size of random occupied-virtual Fock elements (zero for a
             canonical RHF reference).
    """
    rng = np.random.default_rng(seed)
    n = no + nv
    rank = rank or max(4, n)
    B = rng.standard_normal((n, n, rank)) * scale / np.sqrt(rank) ** 0.5
    eri = chemist_from_factors(B)
    ints = dirac_from_chemist(eri)
    eo = -np.sort(rng.uniform(0.4, 2.0, no))[::-1]
    ev = np.sort(rng.uniform(gap, gap + 2.5, nv))
    fock = np.diag(np.concatenate([eo, ev]))
    if offdiag:
        X = rng.standard_normal((n, n)) * offdiag
        X = 0.5 * (X + X.T)
        X[:no, no:] = 0.0
        X[no:, :no] = 0.0
        np.fill_diagonal(X, 0.0)
        fock = fock + X
    if fov:
        Y = rng.standard_normal((no, nv)) * fov
        fock[:no, no:] += Y
        fock[no:, :no] += Y.T
    return np.ascontiguousarray(fock), ints


def direct_sum(hA, noA, hB, noB):
    """Hamiltonian of two non-interacting fragments A and B, orbitals ordered
    (occ A, occ B, vir A, vir B).  CCSD and (T) are size extensive, so
    E(AB) = E(A) + E(B)."""
    fA, iA = hA
    fB, iB = hB
    nA, nB = fA.shape[0], fB.shape[0]
    nvA, nvB = nA - noA, nB - noB
    no, n = noA + noB, nA + nB
    mapA = np.concatenate([np.arange(noA), no + np.arange(nvA)])
    mapB = np.concatenate([noA + np.arange(noB), no + nvA + np.arange(nvB)])
    f = np.zeros((n, n))
    f[np.ix_(mapA, mapA)] = fA
    f[np.ix_(mapB, mapB)] = fB
    ints = np.zeros((n, n, n, n))
    ints[np.ix_(mapA, mapA, mapA, mapA)] = iA
    ints[np.ix_(mapB, mapB, mapB, mapB)] = iB
    return f, ints


def rotate(fock, ints, no, seed=0, angle=0.3):
    """Applies random occupied-occupied and virtual-virtual rotations."""
    rng = np.random.default_rng(seed)
    n = fock.shape[0]

    def orth(k):
        A = rng.standard_normal((k, k)) * angle
        A = A - A.T
        w, V = np.linalg.eigh(1j * A)
        return np.real(V @ np.diag(np.exp(-1j * w)) @ V.conj().T)

    U = np.zeros((n, n))
    U[:no, :no] = orth(no)
    U[no:, no:] = orth(n - no)
    f = U.T @ fock @ U
    g = np.einsum("pqrs,pa,qb,rc,sd->abcd", ints, U, U, U, U, optimize=True)
    return f, g


def random_blocks(no, nv, seed=0, scale=0.08, rank=None, gap=0.5, dtype=np.float64):
    """(fock, blocks) for large sizes without forming the full nact^4 array.

    Same construction as random_hamiltonian -- chemist integrals
    (pr|qs) = sum_K B[p,r,K] B[q,s,K] with B symmetric in its orbital pair --
    but each of the six Dirac blocks <pq|rs> = (pr|qs) is formed directly
    from the factor blocks, so memory stays at the size of the blocks."""
    rng = np.random.default_rng(seed)
    n = no + nv
    rank = rank or max(4, n)
    B = rng.standard_normal((n, n, rank)) * scale / np.sqrt(rank) ** 0.5
    B = 0.5 * (B + B.transpose(1, 0, 2))
    o, v = slice(0, no), slice(no, n)

    def block(P, Q, R, S):
        # <pq|rs> = sum_K B[p,r,K] B[q,s,K] -> [p][q][r][s]
        X = B[P, R]  # (p, r, K)
        Y = B[Q, S]  # (q, s, K)
        p_, r_ = X.shape[:2]
        q_, s_ = Y.shape[:2]
        G = X.reshape(p_ * r_, rank) @ Y.reshape(q_ * s_, rank).T  # [(p r), (q s)]
        return np.ascontiguousarray(G.reshape(p_, r_, q_, s_).transpose(0, 2, 1, 3))

    blocks = dict(oooo=block(o, o, o, o), ooov=block(o, o, o, v), oovv=block(o, o, v, v),
                  ovov=block(o, v, o, v), ovvv=block(o, v, v, v), vvvv=block(v, v, v, v))
    eo = -np.sort(rng.uniform(0.4, 2.0, no))[::-1]
    ev = np.sort(rng.uniform(gap, gap + 2.5, nv))
    fock = np.diag(np.concatenate([eo, ev]))
    return np.ascontiguousarray(fock), blocks
