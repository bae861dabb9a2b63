"""Independent pins for the oracle: textbook spin-orbital CCSD and (T), and
brute-force FCI for two electrons.  None of this follows the reference's
spin-adapted equations; agreement with them is what pins the oracle.

Spin-orbital CCSD: Stanton and Gauss, J. Chem. Phys. 94, 4334 (1991)
intermediates (Fae, Fmi, Fme, Wmnij, Wabef, Wmbej) with a Jacobi update on
the diagonal-Fock denominators.  (T): Raghavachari et al., Chem. Phys. Lett.
157, 479 (1989), the connected/disconnected triples form for canonical
orbitals.
"""
import numpy as np


def spin_orbital(fock, ints, no):
    """Spatial (fock, <pq|rs>) -> spin-orbital (f, <PQ||RS>), occupied first.
    Spin orbital P = 2p + sigma."""
    n = fock.shape[0]
    N = 2 * n
    spin = np.arange(N) % 2
    sp = np.arange(N) // 2
    f = fock[np.ix_(sp, sp)] * (spin[:, None] == spin[None, :])
    g = ints[np.ix_(sp, sp, sp, sp)]
    same = spin[:, None] == spin[None, :]
    coul = g * same[:, None, :, None] * same[None, :, None, :]
    exch = g.transpose(0, 1, 3, 2) * same[:, None, None, :] * same[None, :, :, None]
    return f, coul - exch, 2 * no


def ccsd_spinorb(fock, ints, no, conv=1e-12, maxiter=500):
    f, g, O = spin_orbital(fock, ints, no)
    N = f.shape[0]
    o, v = slice(0, O), slice(O, N)
    fo = np.diag(f)[o]
    fv = np.diag(f)[v]
    Dia = fo[:, None] - fv[None, :]
    Dijab = fo[:, None, None, None] + fo[None, :, None, None] - fv[None, None, :, None] - fv[None, None, None, :]
    t1 = f[o, v] / Dia
    t2 = g[o, o, v, v] / Dijab
    fov = f[o, v]
    foo_off = f[o, o] - np.diag(np.diag(f[o, o]))
    fvv_off = f[v, v] - np.diag(np.diag(f[v, v]))

    def energy(t1, t2):
        return (np.einsum("ia,ia", fov, t1) + 0.25 * np.einsum("ijab,ijab", g[o, o, v, v], t2)
                + 0.5 * np.einsum("ijab,ia,jb", g[o, o, v, v], t1, t1))

    eold = energy(t1, t2)
    # DIIS (plain Pulay over the full amplitude vector) to keep tiny systems fast
    hist_t, hist_e = [], []
    for it in range(maxiter):
        tt = t2 + 0.5 * (np.einsum("ia,jb->ijab", t1, t1) - np.einsum("ib,ja->ijab", t1, t1))
        tau = t2 + np.einsum("ia,jb->ijab", t1, t1) - np.einsum("ib,ja->ijab", t1, t1)
        Fae = (fvv_off - 0.5 * np.einsum("me,ma->ae", fov, t1) + np.einsum("mf,mafe->ae", t1, g[o, v, v, v])
               - 0.5 * np.einsum("mnaf,mnef->ae", tt, g[o, o, v, v]))
        Fmi = (foo_off + 0.5 * np.einsum("ie,me->mi", t1, fov) + np.einsum("ne,mnie->mi", t1, g[o, o, o, v])
               + 0.5 * np.einsum("inef,mnef->mi", tt, g[o, o, v, v]))
        Fme = fov + np.einsum("nf,mnef->me", t1, g[o, o, v, v])
        Wmnij = (g[o, o, o, o] + np.einsum("je,mnie->mnij", t1, g[o, o, o, v])
                 - np.einsum("ie,mnje->mnij", t1, g[o, o, o, v])
                 + 0.25 * np.einsum("ijef,mnef->mnij", tau, g[o, o, v, v]))
        Wabef = (g[v, v, v, v] - np.einsum("mb,amef->abef", t1, g[v, o, v, v])
                 + np.einsum("ma,bmef->abef", t1, g[v, o, v, v])
                 + 0.25 * np.einsum("mnab,mnef->abef", tau, g[o, o, v, v]))
        Wmbej = (g[o, v, v, o] + np.einsum("jf,mbef->mbej", t1, g[o, v, v, v])
                 - np.einsum("nb,mnej->mbej", t1, g[o, o, v, o])
                 - np.einsum("jnfb,mnef->mbej", 0.5 * t2 + np.einsum("jf,nb->jnfb", t1, t1), g[o, o, v, v]))
        r1 = (fov + np.einsum("ie,ae->ia", t1, Fae) - np.einsum("ma,mi->ia", t1, Fmi)
              + np.einsum("imae,me->ia", t2, Fme) - np.einsum("nf,naif->ia", t1, g[o, v, o, v])
              - 0.5 * np.einsum("imef,maef->ia", t2, g[o, v, v, v])
              - 0.5 * np.einsum("mnae,nmei->ia", t2, g[o, o, v, o]))
        Ft = Fae - 0.5 * np.einsum("mb,me->be", t1, Fme)
        Gt = Fmi + 0.5 * np.einsum("je,me->mj", t1, Fme)
        X = np.einsum("ijae,be->ijab", t2, Ft)
        r2 = g[o, o, v, v] + X - X.transpose(0, 1, 3, 2)
        Y = np.einsum("imab,mj->ijab", t2, Gt)
        r2 -= Y - Y.transpose(1, 0, 2, 3)
        r2 += 0.5 * np.einsum("mnab,mnij->ijab", tau, Wmnij)
        r2 += 0.5 * np.einsum("ijef,abef->ijab", tau, Wabef)
        Z = np.einsum("imae,mbej->ijab", t2, Wmbej) - np.einsum("ie,ma,mbej->ijab", t1, t1, g[o, v, v, o])
        r2 += Z - Z.transpose(1, 0, 2, 3) - Z.transpose(0, 1, 3, 2) + Z.transpose(1, 0, 3, 2)
        Q = np.einsum("ie,abej->ijab", t1, g[v, v, v, o])
        r2 += Q - Q.transpose(1, 0, 2, 3)
        P = np.einsum("ma,mbij->ijab", t1, g[o, v, o, o])
        r2 -= P - P.transpose(0, 1, 3, 2)
        t1n, t2n = r1 / Dia, r2 / Dijab
        # residual in the "t D = rhs" form: rhs/D - t
        err = np.concatenate([(t1n - t1).ravel(), (t2n - t2).ravel()])
        hist_t.append(np.concatenate([t1n.ravel(), t2n.ravel()]))
        hist_e.append(err)
        if len(hist_t) > 8:
            hist_t.pop(0); hist_e.pop(0)
        k = len(hist_t)
        if k >= 3:
            B = -np.ones((k + 1, k + 1)); B[k, k] = 0
            E = np.array(hist_e)
            B[:k, :k] = E @ E.T
            rhs = np.zeros(k + 1); rhs[k] = -1
            c = np.linalg.solve(B, rhs)[:k]
            vec = c @ np.array(hist_t)
        else:
            vec = hist_t[-1]
        t1 = vec[: t1.size].reshape(t1.shape)
        t2 = vec[t1.size:].reshape(t2.shape)
        e = energy(t1, t2)
        if abs(e - eold) < conv and np.linalg.norm(err) < conv * 10:
            break
        eold = e
    return e, t1, t2, (f, g, O)


def triples_spinorb(fock, ints, no, conv=1e-12):
    e, t1, t2, (f, g, O) = ccsd_spinorb(fock, ints, no, conv=conv)
    N = f.shape[0]
    o, v = slice(0, O), slice(O, N)
    fo = np.diag(f)[o]
    fv = np.diag(f)[v]
    D3 = (fo[:, None, None, None, None, None] + fo[None, :, None, None, None, None] + fo[None, None, :, None, None, None]
          - fv[None, None, None, :, None, None] - fv[None, None, None, None, :, None] - fv[None, None, None, None, None, :])

    def P_i_jk(x):  # P(i/jk) on axes 0,1,2
        return x - x.transpose(1, 0, 2, 3, 4, 5) - x.transpose(2, 1, 0, 3, 4, 5)

    def P_a_bc(x):
        return x - x.transpose(0, 1, 2, 4, 3, 5) - x.transpose(0, 1, 2, 5, 4, 3)

    disc = np.einsum("ia,jkbc->ijkabc", t1, g[o, o, v, v])
    conn = np.einsum("jkae,eibc->ijkabc", t2, g[v, o, v, v]) - np.einsum("imbc,majk->ijkabc", t2, g[o, v, o, o])
    td = P_i_jk(P_a_bc(disc)) / D3
    tc = P_i_jk(P_a_bc(conn)) / D3
    et = np.einsum("ijkabc,ijkabc", tc * D3, tc + td) / 36.0
    return e, et


def fci_two_electron(fock, ints):
    """Lowest singlet energy of a 2-electron system minus the reference energy
    of the doubly occupied orbital 0 (the exact correlation energy)."""
    n = fock.shape[0]
    J = np.einsum("pqq->p", ints[:, 0, :, 0][:, :, None] * np.eye(n)[None]) if False else None
    h = fock - 2.0 * ints[:, 0, :, 0] + ints[:, 0, 0, :]
    eref = 2 * h[0, 0] + ints[0, 0, 0, 0]
    I = np.eye(n)
    H = (np.einsum("pr,qs->pqrs", h, I) + np.einsum("pr,qs->pqrs", I, h) + ints).reshape(n * n, n * n)
    # orthonormal basis of symmetric two-orbital functions
    cols = []
    for p in range(n):
        for q in range(p, n):
            c = np.zeros((n, n))
            if p == q:
                c[p, p] = 1.0
            else:
                c[p, q] = c[q, p] = 1 / np.sqrt(2)
            cols.append(c.ravel())
    S = np.array(cols).T
    Hs = S.T @ H @ S
    return np.linalg.eigvalsh(0.5 * (Hs + Hs.T))[0] - eref
