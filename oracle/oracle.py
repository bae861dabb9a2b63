"""ctypes binding of ugacc_oracle.c (test infrastructure only).

Each method names the reference routine it restates; the C source carries
the file:line citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def libpath(omp=False):
    return os.path.join(_HERE, "liboracle_omp.so" if omp else "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def _lib(omp=False):
    if omp not in _libs:
        path = libpath(omp)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_destroy.argtypes = [vp]
        for fn in ("orc_energy", "orc_energy_full", "orc_t1norm", "orc_tcorr", "orc_tcorr_ooc", "orc_tcorr_tjl"):
            getattr(L, fn).restype = ctypes.c_double
            getattr(L, fn).argtypes = [vp]
        L.orc_tcorr_tjl_range.restype = ctypes.c_double
        L.orc_tcorr_tjl_range.argtypes = [vp, ctypes.c_long, ctypes.c_long]
        L.orc_tjl_fov_range.restype = ctypes.c_double
        L.orc_tjl_fov_range.argtypes = [vp, ctypes.c_long, ctypes.c_long]
        L.orc_ntriplets.restype = ctypes.c_long
        L.orc_ntriplets.argtypes = [ctypes.c_int]
        L.orc_set_amps.argtypes = [vp, dp, dp]
        L.orc_get_amps.argtypes = [vp, dp, dp]
        L.orc_get_tau.argtypes = [vp, dp, dp]
        L.orc_get_L.argtypes = [vp, dp]
        L.orc_build_tstar.argtypes = [vp]
        L.orc_iterate.restype = ctypes.c_double
        L.orc_iterate.argtypes = [vp, dp, dp]
        L.orc_get_intermediates.argtypes = [vp] + [dp] * 7
        L.orc_ccsd.restype = ctypes.c_int
        L.orc_ccsd.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, dp, dp]
        L.orc_tjl_per_triplet.argtypes = [vp, dp]
        L.orc_tjl_fov_per_triplet.argtypes = [vp, dp]
        L.orc_tjl_fov_list.argtypes = [vp, ctypes.c_long, ctypes.POINTER(ctypes.c_long), dp]
        L.orc_w3.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.orc_iteration_sample.argtypes = [vp, ctypes.c_int]
        L.orc_hbar.argtypes = [vp] + [dp] * 11
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_residual_rows.argtypes = [vp, ctypes.c_int, dp, dp]
        L.orc_lambda_init.argtypes = [vp]
        L.orc_lambda_residual.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.orc_lambda.restype = ctypes.c_int
        L.orc_lambda.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_get_lambda.argtypes = [vp, dp, dp]
        L.orc_set_lambda.argtypes = [vp, dp, dp]
        L.orc_lambda_pseudoenergy.restype = ctypes.c_double
        L.orc_density.argtypes = [vp] + [dp] * 11
        L.orc_lambda_pseudoenergy.argtypes = [vp]
        L.orc_tgrad.argtypes = [vp] + [dp] * 8
        L.orc_tgrad_ooc.argtypes = [vp] + [dp] * 8
        L.orc_pert_create.restype = vp
        L.orc_pert_create.argtypes = [vp, dp, ctypes.c_double]
        L.orc_pert_destroy.argtypes = [vp]
        L.orc_pert_solve.restype = ctypes.c_int
        L.orc_pert_solve.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp]
        L.orc_pert_pseudoresponse.restype = ctypes.c_double
        L.orc_pert_pseudoresponse.argtypes = [vp, ctypes.c_int]
        L.orc_pert_residual.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp]
        L.orc_pert_get.argtypes = [vp, dp, dp, dp, dp]
        L.orc_pert_set.argtypes = [vp, dp, dp, dp, dp]
        L.orc_linresp.restype = ctypes.c_double
        L.orc_linresp.argtypes = [vp, vp]
        L.orc_tgrad_enable.argtypes = [vp, ctypes.c_int]
        _libs[omp] = L
    return _libs[omp]


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """State of one CCWfn (ccwfn.cc:12-120) over a Hamiltonian (fock, ints)."""

    def __init__(self, fock, ints, no, omp=False):
        self.L = _lib(omp)
        self.no = no
        self.nv = fock.shape[0] - no
        self._fock = np.ascontiguousarray(fock, dtype=np.float64)
        self._ints = np.ascontiguousarray(ints, dtype=np.float64)
        self.h = self.L.orc_create(no, self.nv, _p(self._fock), _p(self._ints))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def threads(self):
        return self.L.orc_num_threads()

    def set_threads(self, n):
        """OpenMP threads of this library's loops (process-wide)."""
        self.L.orc_set_threads(int(n))

    def energy(self):
        return self.L.orc_energy(self.h)

    def energy_full(self):
        return self.L.orc_energy_full(self.h)

    def t1norm(self):
        return self.L.orc_t1norm(self.h)

    def set_amps(self, t1, t2):
        t1 = np.ascontiguousarray(t1, dtype=np.float64)
        t2 = np.ascontiguousarray(t2, dtype=np.float64)
        self.L.orc_set_amps(self.h, _p(t1), _p(t2))

    def amps(self):
        t1 = np.zeros((self.no, self.nv))
        t2 = np.zeros((self.no, self.no, self.nv, self.nv))
        self.L.orc_get_amps(self.h, _p(t1), _p(t2))
        return t1, t2

    def tau(self):
        tau = np.zeros((self.no, self.no, self.nv, self.nv))
        ttau = np.zeros_like(tau)
        self.L.orc_get_tau(self.h, _p(tau), _p(ttau))
        return tau, ttau

    def Lints(self):
        n = self.no + self.nv
        L = np.zeros((n, n, n, n))
        self.L.orc_get_L(self.h, _p(L))
        return L

    def iterate(self):
        """One Jacobi step; returns (rms, R1, R2) with the raw residuals."""
        R1 = np.zeros((self.no, self.nv))
        R2 = np.zeros((self.no, self.no, self.nv, self.nv))
        rms = self.L.orc_iterate(self.h, _p(R1), _p(R2))
        return rms, R1, R2

    def intermediates(self):
        o, v = self.no, self.nv
        out = dict(Fvv=np.zeros((v, v)), Foo=np.zeros((o, o)), Fov=np.zeros((o, v)),
                   Woooo=np.zeros((o, o, o, o)), Wovov=np.zeros((o, v, o, v)),
                   Wovvo=np.zeros((o, v, v, o)), Wovoo=np.zeros((o, v, o, o)))
        self.L.orc_get_intermediates(self.h, *[_p(out[k]) for k in
                                               ("Fvv", "Foo", "Fov", "Woooo", "Wovov", "Wovvo", "Wovoo")])
        return out

    def ccsd(self, maxiter=100, conv=1e-7, diis=True, mode=0, diis_policy=0, min_vecs=3):
        """compute_energy's iteration (ccwfn.cc:188-231) for mode 0, the
        attic driver (attic/ugacc.cc:100-114) for mode 1."""
        out = np.zeros(4)
        trace = np.zeros(maxiter + 1)
        rc = self.L.orc_ccsd(self.h, maxiter, conv, int(diis), mode, diis_policy, min_vecs, _p(out), _p(trace))
        n = int(out[3])
        return dict(converged=(rc == 0), emp2=out[0], eccsd=out[1], rms=out[2], niter=n,
                    trace=trace[: n + 1].copy())

    def build_tstar(self):
        self.L.orc_build_tstar(self.h)

    def tcorr(self):
        return self.L.orc_tcorr(self.h)

    def tcorr_ooc(self):
        return self.L.orc_tcorr_ooc(self.h)

    def tcorr_tjl(self):
        return self.L.orc_tcorr_tjl(self.h)

    def tcorr_tjl_range(self, t0, t1):
        return self.L.orc_tcorr_tjl_range(self.h, t0, t1)

    def tjl_fov_range(self, t0, t1):
        """Triplets [t0, t1) of the TJL bracket with tcorr's f_ov terms
        folded in (the full range is tcorr's value)."""
        return self.L.orc_tjl_fov_range(self.h, t0, t1)

    def ntriplets(self):
        return self.L.orc_ntriplets(self.no)

    def tjl_per_triplet(self):
        e = np.zeros(self.ntriplets())
        self.L.orc_tjl_per_triplet(self.h, _p(e))
        return e

    def tjl_fov_per_triplet(self):
        """Per-triplet TJL energies with tcorr's f_ov t3 terms folded into V
        (their sum is tcorr's value for any f_ov)."""
        e = np.zeros(self.ntriplets())
        self.L.orc_tjl_fov_per_triplet(self.h, _p(e))
        return e

    def tjl_fov_list(self, idx):
        """Folded per-triplet energies of the triplets idx (i>=j>=k order)."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        e = np.zeros(idx.size)
        self.L.orc_tjl_fov_list(self.h, idx.size, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), _p(e))
        return e

    def w3(self, i, j, k):
        W = np.zeros((self.nv,) * 3)
        self.L.orc_w3(self.h, i, j, k, _p(W))
        return W

    HBAR_BLOCKS = ("Hov", "Hoo", "Hvv", "Hoooo", "Hvvvv", "Hvovv", "Hooov", "Hovvo", "Hovov", "Hvvvo", "Hovoo")

    @staticmethod
    def hbar_shapes(no, nv):
        o, v = no, nv
        return dict(Hov=(o, v), Hoo=(o, o), Hvv=(v, v), Hoooo=(o, o, o, o), Hvvvv=(v, v, v, v), Hvovv=(v, o, v, v),
                    Hooov=(o, o, o, v), Hovvo=(o, v, v, o), Hovov=(o, v, o, v), Hvvvo=(v, v, v, o),
                    Hovoo=(o, v, o, o))

    def hbar(self):
        """HBAR::HBAR (hbar.cc:8-228) from the current amplitudes: the eleven
        blocks by name, in the reference's layouts."""
        sh = self.hbar_shapes(self.no, self.nv)
        out = {k: np.zeros(sh[k]) for k in self.HBAR_BLOCKS}
        self.L.orc_hbar(self.h, *[_p(out[k]) for k in self.HBAR_BLOCKS])
        return out

    def lambda_init(self):
        """CCLambda::CCLambda (cclambda.cc:13-46): HBAR of the current
        amplitudes and the guess l1 = 2 t1, l2 = 2 (2 t2 - t2^T)."""
        self.L.orc_lambda_init(self.h)

    def lambda_residual(self, l1, l2):
        """build_G, build_l1, build_l2 (cclambda.cc:137-268) at (l1, l2) over
        the HBAR of lambda_init: (R1, R2, Goo, Gvv), R raw (before /D)."""
        no, nv = self.no, self.nv
        l1 = np.ascontiguousarray(l1, dtype=np.float64)
        l2 = np.ascontiguousarray(l2, dtype=np.float64)
        R1, R2 = np.zeros((no, nv)), np.zeros((no, no, nv, nv))
        Goo, Gvv = np.zeros((no, no)), np.zeros((nv, nv))
        self.L.orc_lambda_residual(self.h, _p(l1), _p(l2), _p(R1), _p(R2), _p(Goo), _p(Gvv))
        return R1, R2, Goo, Gvv

    def lambda_solve(self, maxiter=100, conv=1e-7, diis=True, diis_policy=0, min_vecs=3):
        """CCLambda::compute_lambda (cclambda.cc:93-135) from the current
        guess (lambda_init first)."""
        out = np.zeros(3)
        trace = np.zeros(maxiter + 1)
        rc = self.L.orc_lambda(self.h, maxiter, conv, int(diis), diis_policy, min_vecs, _p(out), _p(trace))
        n = int(out[2])
        return dict(converged=rc == 0, pseudoenergy=out[0], rms=out[1], niter=n, trace=trace[:n + 1])

    def lambdas(self):
        l1 = np.zeros((self.no, self.nv))
        l2 = np.zeros((self.no, self.no, self.nv, self.nv))
        self.L.orc_get_lambda(self.h, _p(l1), _p(l2))
        return l1, l2

    def set_lambdas(self, l1, l2):
        l1 = np.ascontiguousarray(l1, dtype=np.float64)
        l2 = np.ascontiguousarray(l2, dtype=np.float64)
        self.L.orc_set_lambda(self.h, _p(l1), _p(l2))

    def pseudoenergy(self):
        """cclambda.cc:270-288 of the current l2."""
        return self.L.orc_lambda_pseudoenergy(self.h)

    DENSITY_BLOCKS = ("Doo", "Dvv", "Dov", "Dvo", "Goooo", "Gvvvv", "Gooov", "Gvvvo", "Govov", "Goovv")
    DENSITY_PARTS = ("oooo", "vvvv", "ooov", "vvvo", "ovov", "oovv")

    def density(self):
        """CCDensity::onepdm + twopdm (ccdensity.cc:89-339) from the current
        amplitudes and the Lambda amplitudes: e_one, e_two, the six parts of
        e_two by name and the ten density blocks."""
        o, v = self.no, self.nv
        sh = dict(Doo=(o, o), Dvv=(v, v), Dov=(o, v), Dvo=(v, o), Goooo=(o, o, o, o), Gvvvv=(v, v, v, v),
                  Gooov=(o, o, o, v), Gvvvo=(v, v, v, o), Govov=(o, v, o, v), Goovv=(o, o, v, v))
        blocks = {k: np.zeros(sh[k]) for k in self.DENSITY_BLOCKS}
        E = np.zeros(8)
        self.L.orc_density(self.h, _p(E), *[_p(blocks[k]) for k in self.DENSITY_BLOCKS])
        return dict(e_one=E[0], e_two=E[1], parts=dict(zip(self.DENSITY_PARTS, E[2:])), blocks=blocks)

    TGRAD_BLOCKS = ("s1", "s2", "Doo", "Dvv", "Dov", "Gooov", "Gvvvo", "Goovv")

    def tgrad(self, enable=True):
        """CCWfn::tgrad (ccwfn.cc:854-1098) from the current amplitudes: the
        Lambda sources s1, s2 and the (T) density parts by name.  enable: the
        following lambda_solve / lambda_residual / density include them
        (cclambda.cc:181, :243; ccdensity.cc:43-74)."""
        o, v = self.no, self.nv
        sh = dict(s1=(o, v), s2=(o, o, v, v), Doo=(o, o), Dvv=(v, v), Dov=(o, v), Gooov=(o, o, o, v),
                  Gvvvo=(v, v, v, o), Goovv=(o, o, v, v))
        out = {k: np.zeros(sh[k]) for k in self.TGRAD_BLOCKS}
        self.L.orc_tgrad(self.h, *[_p(out[k]) for k in self.TGRAD_BLOCKS])
        self.L.orc_tgrad_enable(self.h, int(enable))
        return out

    def tgrad_ooc(self, enable=True):
        """CCWfn::tgrad_ooc (ccwfn.cc:1150-1316), the OOC=true form: the same
        blocks as tgrad(), replacing its terms in the state."""
        o, v = self.no, self.nv
        sh = dict(s1=(o, v), s2=(o, o, v, v), Doo=(o, o), Dvv=(v, v), Dov=(o, v), Gooov=(o, o, o, v),
                  Gvvvo=(v, v, v, o), Goovv=(o, o, v, v))
        out = {k: np.zeros(sh[k]) for k in self.TGRAD_BLOCKS}
        self.L.orc_tgrad_ooc(self.h, *[_p(out[k]) for k in self.TGRAD_BLOCKS])
        self.L.orc_tgrad_enable(self.h, int(enable))
        return out

    def tgrad_enable(self, on):
        self.L.orc_tgrad_enable(self.h, int(on))

    def pert(self, A, omega=0.0):
        """CCPert::CCPert (ccpert.cc:11-75) of the one-electron perturbation A
        (nact x nact) at omega, over the HBAR of lambda_init and the current
        Lambda amplitudes."""
        return Pert(self, A, omega)

    def residual_rows(self, rows):
        """(R1, R2[:rows]) of the current amplitudes."""
        R1 = np.zeros((self.no, self.nv))
        R2 = np.zeros((rows, self.no, self.nv, self.nv))
        self.L.orc_residual_rows(self.h, rows, _p(R1), _p(R2))
        return R1, R2

    def iteration_sample(self, rows):
        """Runs rows/no of one CCSD iteration (for the bounded CPU baseline)."""
        self.L.orc_iteration_sample(self.h, rows)


class Pert:
    """One perturbed wave function (ccpert.cc): pertbar, the right-hand (X)
    and left-hand (Y) perturbed amplitudes, and CCResp::linresp."""

    def __init__(self, orc, A, omega):
        self.o = orc
        self.L = orc.L
        self.A = np.ascontiguousarray(A, dtype=np.float64)
        self.omega = omega
        self.h = self.L.orc_pert_create(orc.h, _p(self.A), omega)

    def close(self):
        if self.h:
            self.L.orc_pert_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, hand=0, maxiter=200, conv=1e-10, diis=True, min_vecs=3):
        """CCPert::solve (ccpert.cc:178-225): hand 0 right (X), 1 left (Y)."""
        out = np.zeros(3)
        rc = self.L.orc_pert_solve(self.h, hand, maxiter, conv, int(diis), min_vecs, _p(out))
        return dict(converged=rc == 0, pseudoresponse=out[0], rms=out[1], niter=int(out[2]))

    def pseudoresponse(self, hand=0):
        return self.L.orc_pert_pseudoresponse(self.h, hand)

    def residual(self, hand, A1, A2):
        no, nv = self.o.no, self.o.nv
        A1 = np.ascontiguousarray(A1, dtype=np.float64)
        A2 = np.ascontiguousarray(A2, dtype=np.float64)
        R1, R2 = np.zeros((no, nv)), np.zeros((no, no, nv, nv))
        self.L.orc_pert_residual(self.h, hand, _p(A1), _p(A2), _p(R1), _p(R2))
        return R1, R2

    def amplitudes(self):
        no, nv = self.o.no, self.o.nv
        X1, Y1 = np.zeros((no, nv)), np.zeros((no, nv))
        X2, Y2 = np.zeros((no, no, nv, nv)), np.zeros((no, no, nv, nv))
        self.L.orc_pert_get(self.h, _p(X1), _p(X2), _p(Y1), _p(Y2))
        return X1, X2, Y1, Y2

    def set_amplitudes(self, X1=None, X2=None, Y1=None, Y2=None):
        c = lambda a: None if a is None else _p(np.ascontiguousarray(a, dtype=np.float64))
        keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None for a in (X1, X2, Y1, Y2)]
        self.L.orc_pert_set(self.h, *[None if a is None else _p(a) for a in keep])

    def linresp(self, other):
        """CCResp::linresp (ccresp.cc:27-124): <<A;B>> with A this
        perturbation and B other's."""
        return self.L.orc_linresp(self.h, other.h)
