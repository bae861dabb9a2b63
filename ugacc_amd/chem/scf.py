"""Closed-shell RHF and the MO-basis Hamiltonian the CC code consumes.

This is the front end that stands in for the Psi4 wavefunction the
reference plugin receives (ugacc.cc:75-99: a converged RHF reference) and
for Hamiltonian's construction of the MO Fock matrix and integrals
(hamiltonian.cc:19-103): QT orbital order (frozen core, occupied, virtual by
orbital energy in C1), the active-space Fock matrix taken from the SCF Fock
(hamiltonian.cc:47-56), the frozen-core energy, and the six Dirac blocks
<pq|rs> = (pr|qs) over active orbitals.
"""
import numpy as np

from . import integrals
from .basis import Z as ZTAB
from .basis import shells_for


def nuclear_repulsion(atoms):
    e = 0.0
    for i in range(len(atoms)):
        for j in range(i):
            zi, zj = ZTAB[atoms[i][0]], ZTAB[atoms[j][0]]
            d = np.linalg.norm(np.subtract(atoms[i][1], atoms[j][1]))
            e += zi * zj / d
    return e


class RHF:
    """Converged closed-shell RHF (DIIS on FDS - SDF, core-Hamiltonian guess)."""

    def __init__(self, atoms, basis, charge=0, conv=1e-12, maxiter=200, eri_thresh=1e-14, log=None):
        self.atoms = atoms
        self.basis = basis
        shells = shells_for(atoms, basis)
        self.shells = shells
        Zs = [ZTAB[a] for a, _ in atoms]
        nel = sum(Zs) - charge
        if nel % 2:
            raise ValueError("closed-shell RHF needs an even electron count")
        self.nocc = nel // 2
        self.enuc = nuclear_repulsion(atoms)
        S, T, V = integrals.one_electron(shells, Zs, [x for _, x in atoms])
        self.S, self.H = S, T + V
        if log:
            log("AO integrals: %d basis functions" % S.shape[0])
        self.eri = integrals.eri(shells, eri_thresh)
        self.nbf = S.shape[0]
        self._solve(conv, maxiter, log)

    def _solve(self, conv, maxiter, log):
        S, H, no = self.S, self.H, self.nocc
        w, U = np.linalg.eigh(S)
        X = U @ np.diag(w ** -0.5) @ U.T
        F = H.copy()
        Fs, Es = [], []
        e_old = 0.0
        for it in range(1, maxiter + 1):
            e, Cp = np.linalg.eigh(X.T @ F @ X)
            C = X @ Cp
            D = C[:, :no] @ C[:, :no].T
            J, K = integrals.jk(self.eri, D)
            F = H + 2 * J - K
            energy = np.sum(D * (H + F)) + self.enuc
            err = X.T @ (F @ D @ S - S @ D @ F) @ X
            Fs.append(F.copy())
            Es.append(err)
            if len(Fs) > 8:
                Fs.pop(0)
                Es.pop(0)
            dnorm = np.max(np.abs(err))
            if log:
                log("SCF iter %2d  E = %.12f  dE = %.3e  |[F,D]| = %.3e" % (it, energy, energy - e_old, dnorm))
            if abs(energy - e_old) < conv and dnorm < conv * 10:
                break
            e_old = energy
            if len(Fs) >= 2:
                k = len(Fs)
                B = -np.ones((k + 1, k + 1))
                B[k, k] = 0
                for a in range(k):
                    for b in range(k):
                        B[a, b] = np.sum(Es[a] * Es[b])
                rhs = np.zeros(k + 1)
                rhs[k] = -1
                c = np.linalg.solve(B, rhs)[:k]
                F = sum(ci * Fi for ci, Fi in zip(c, Fs))
        else:
            raise RuntimeError("SCF has not converged")
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        self.C = X @ Cp
        self.eps = e
        self.F_ao = F
        self.energy = energy
        self.iterations = it


def transform_blocks(eri_ao, C, nfzc, nocc, device=None):
    """Six Dirac blocks over active orbitals: <pq|rs> = (pr|qs), with the
    chemist integrals (pr|qs) transformed from the AO basis as GEMMs: the ket
    pair first (one n^3 x n GEMM and a batched n x n one), then the bra pair,
    each half-transformed ket block shared by the blocks that need it.

    The GEMMs run through torch (on the GPU when one is present, else on the
    host cores); this is set-up work outside the CC solve."""
    import torch

    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    dt = torch.float64
    n = eri_ao.shape[0]
    g = torch.from_numpy(np.ascontiguousarray(eri_ao)).to(device)
    Co = torch.from_numpy(np.ascontiguousarray(C[:, nfzc:nocc])).to(device, dt)
    Cv = torch.from_numpy(np.ascontiguousarray(C[:, nocc:])).to(device, dt)
    flat = g.reshape(n * n * n, n)

    def ket(C3, C4):
        # H[p q][k l] = sum_rs (pq|rs) C3[r,k] C4[s,l]
        A = (flat @ C4).reshape(n * n, n, C4.shape[1])              # (pq) r l
        H = torch.matmul(C3.T.unsqueeze(0), A)                        # (pq) k l
        return H.reshape(n, n, C3.shape[1] * C4.shape[1])

    def bra(H, C1, C2):
        # X[i][j (kl)] = sum_pq C1[p,i] C2[q,j] H[p q][k l]
        Y = torch.matmul(C2.T.unsqueeze(0), H)                        # p j (kl)
        return C1.T @ Y.reshape(n, -1)                                # i (j kl)

    def dirac(X, n1, n2, n3, n4):
        # chemist (13|24) stored as X[1][3][2][4] -> Dirac <12|34>
        return np.ascontiguousarray(X.reshape(n1, n3, n2, n4).permute(0, 2, 1, 3).contiguous().cpu().numpy())

    o, v = Co.shape[1], Cv.shape[1]
    H = ket(Co, Co)                                   # ket (n j)
    oooo = dirac(bra(H, Co, Co), o, o, o, o)          # <mn|ij> = (mi|nj)
    del H
    H = ket(Co, Cv)                                   # ket (n e) / (n f)
    ooov = dirac(bra(H, Co, Co), o, o, o, v)          # <mn|ie> = (mi|ne)
    oovv = dirac(bra(H, Co, Cv), o, o, v, v)          # <mn|ef> = (me|nf)
    del H
    H = ket(Cv, Cv)                                   # ket (b e) / (b f)
    ovov = dirac(bra(H, Co, Co), o, v, o, v)          # <mb|je> = (mj|be)
    ovvv = dirac(bra(H, Co, Cv), o, v, v, v)          # <mb|ef> = (me|bf)
    vvvv = dirac(bra(H, Cv, Cv), v, v, v, v)          # <ab|ef> = (ae|bf)
    del H, g, flat
    if device != "cpu":
        torch.cuda.empty_cache()
    return dict(oooo=oooo, ooov=ooov, oovv=oovv, ovov=ovov, ovvv=ovvv, vvvv=vvvv)


class Reference:
    """What the plugin entry receives: a converged RHF wavefunction and the
    active-space Hamiltonian built from it (hamiltonian.cc:19-103)."""

    def __init__(self, atoms, basis, nfzc=0, scf_conv=1e-12, log=None):
        self.scf = RHF(atoms, basis, conv=scf_conv, log=log)
        s = self.scf
        self.nfzc = nfzc
        self.no = s.nocc - nfzc
        self.nv = s.nbf - s.nocc
        self.nmo = s.nbf
        C = s.C
        # MO Fock from the SCF Fock, restricted to the active orbitals
        Fmo = C.T @ s.F_ao @ C
        act = slice(nfzc, s.nbf)
        self.fock = np.ascontiguousarray(Fmo[act, act])
        self.eps = s.eps
        # frozen-core energy: sum_c (h + f^core)_cc with the frozen-core Fock
        # operator f^core = h + sum_c' (2 J_c' - K_c')
        self.efzc = 0.0
        if nfzc:
            Dc = C[:, :nfzc] @ C[:, :nfzc].T
            Jc, Kc = integrals.jk(s.eri, Dc)
            self.efzc = float(np.sum(Dc * (2 * s.H + 2 * Jc - Kc)))
        self.escf = s.energy
        self.enuc = s.enuc
        self.blocks = transform_blocks(s.eri, C, nfzc, s.nocc)

    def perturbation(self, op, origin=(0.0, 0.0, 0.0)):
        """The three Cartesian components of a one-electron property operator
        in the active MO basis (nact x nact each), as Perturbation::prop_p
        holds them (perturbation.cc): "Mu" the dipole operator with the
        electron charge, -(r - O) (Psi4's dipole integrals), "L" the magnetic
        dipole -1/2 (r - O) x grad (the angular-momentum integrals without
        their factor -i, scaled by -1/2 as perturbation.cc does)."""
        D, M = integrals.properties(self.scf.shells, origin)
        if op == "Mu":
            X = -D
        elif op == "L":
            X = -0.5 * M
        else:
            raise ValueError("unknown property operator %r (Mu or L)" % op)
        C = self.scf.C[:, self.nfzc:]
        return [np.ascontiguousarray(C.T @ X[k] @ C) for k in range(3)]

    def full_ints(self):
        """nact^4 Dirac array (the reference's ints_ layout, for the oracle)."""
        from ..hamiltonian import full_from_blocks
        return full_from_blocks(self.no, self.nv, self.blocks)
