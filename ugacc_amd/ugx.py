"""ctypes binding of the C ABI in include/ugx.h (libugx.so, built in-tree).

The library is the product path: there is no CPU fallback.  Loading fails
with an ImportError-style message when libugx.so has not been built, and
every call that returns an error status raises ``UgxError`` carrying the
library's own message.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libugx.so")

OK, NOT_CONVERGED, BAD_INPUT, OOM, HIP_ERROR, NO_DEVICE, NOT_READY, NONFINITE = range(8)
PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_BF16X3_PAIR, PREC_BF16, PREC_BF16X3_PANEL = 0, 1, 2, 3, 4, 5
GUESS_MP2, GUESS_USER, GUESS_MLP = 0, 1, 2
POLISH_PLAIN, POLISH_REFINE = 0, 1
MLP_NFEAT, MLP_WIDTH = 8, 128
ENERGY_REFERENCE, ENERGY_FULL = 0, 1
PREC_NAMES = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}

# Every function include/ugx.h declares, in declaration order.
EXPORTS = [
    "ugx_default_options", "ugx_create", "ugx_destroy", "ugx_last_error", "ugx_load_hamiltonian",
    "ugx_set_amplitudes", "ugx_prepare", "ugx_ccsd", "ugx_triples", "ugx_triples_per_triplet", "ugx_w3",
    "ugx_get_amplitudes", "ugx_energy", "ugx_residual", "ugx_intermediates", "ugx_hbar", "ugx_lambda",
    "ugx_get_lambda", "ugx_lambda_residual", "ugx_tgrad", "ugx_tgrad_ooc", "ugx_density", "ugx_pert_create", "ugx_pert_solve",
    "ugx_pert_amplitudes", "ugx_pert_residual", "ugx_linresp", "ugx_pert_release", "ugx_set_profiling",
    "ugx_kernel_stats", "ugx_gemm", "ugx_gemm_bf16x3_static", "ugx_synchronize", "ugx_batch_solve", "ugx_set_autotune", "ugx_tile_table",
    "ugx_save_tile_table", "ugx_load_tile_table", "ugx_mlp_num_params", "ugx_mlp_create", "ugx_mlp_destroy",
    "ugx_mlp_last_error", "ugx_mlp_set_params", "ugx_mlp_get_params", "ugx_mlp_set_scaling", "ugx_mlp_forward",
    "ugx_mlp_loss_grad", "ugx_mlp_train", "ugx_features", "ugx_set_guess_network",
]


class UgxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [
        ("wfn_ccsd_t", ctypes.c_int), ("maxiter", ctypes.c_int), ("diis", ctypes.c_int),
        ("r_convergence", ctypes.c_double), ("ooc", ctypes.c_int), ("diis_nvec", ctypes.c_int),
        ("diis_min_vecs", ctypes.c_int), ("diis_policy", ctypes.c_int), ("precision", ctypes.c_int),
        ("polish_threshold", ctypes.c_double), ("guess", ctypes.c_int), ("energy_form", ctypes.c_int),
        ("print", ctypes.c_int), ("polish_keep_diis", ctypes.c_int), ("graph", ctypes.c_int),
        ("polish_mode", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("e_mp2", ctypes.c_double), ("e_ccsd", ctypes.c_double), ("e_ccsd_final", ctypes.c_double),
        ("e_t", ctypes.c_double), ("rms", ctypes.c_double), ("niter", ctypes.c_int), ("niter_low", ctypes.c_int),
        ("seconds_ccsd", ctypes.c_double), ("seconds_triples", ctypes.c_double), ("flops_ccsd", ctypes.c_double),
        ("device_allocs", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Problem(ctypes.Structure):
    _fields_ = [("no", ctypes.c_int), ("nv", ctypes.c_int)] + [
        (k, ctypes.POINTER(ctypes.c_double)) for k in ("fock", "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")]


_lib = None


def load_library(path=LIBPATH):
    """Loads libugx.so and declares the ABI.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libugx.so is not built (%s); run `python -m ugacc_amd.build`" % path)
    L = ctypes.CDLL(path)
    vp, ip, dp, cd = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double
    L.ugx_default_options.argtypes = [ctypes.POINTER(Options)]
    L.ugx_default_options.restype = None
    L.ugx_create.argtypes = [ctypes.POINTER(vp), ip]
    L.ugx_destroy.argtypes = [vp]
    L.ugx_destroy.restype = None
    L.ugx_last_error.argtypes = [vp]
    L.ugx_last_error.restype = ctypes.c_char_p
    L.ugx_load_hamiltonian.argtypes = [vp, ip, ip] + [dp] * 7
    L.ugx_set_amplitudes.argtypes = [vp, dp, dp]
    L.ugx_prepare.argtypes = [vp, ctypes.POINTER(Options)]
    L.ugx_ccsd.argtypes = [vp, ctypes.POINTER(Options), ctypes.POINTER(Result)]
    L.ugx_triples.argtypes = [vp, ip, dp]
    L.ugx_triples_per_triplet.argtypes = [vp, ip, dp]
    L.ugx_w3.argtypes = [vp, ip, ip, ip, ip, dp]
    L.ugx_get_amplitudes.argtypes = [vp, dp, dp]
    L.ugx_energy.argtypes = [vp, ip, dp]
    L.ugx_residual.argtypes = [vp, dp, dp, ip, dp, dp]
    L.ugx_intermediates.argtypes = [vp] + [dp] * 7
    L.ugx_hbar.argtypes = [vp] + [dp] * 11
    L.ugx_lambda.argtypes = [vp, ctypes.POINTER(Options), dp, ctypes.POINTER(ctypes.c_int), dp, dp]
    L.ugx_get_lambda.argtypes = [vp, dp, dp]
    L.ugx_lambda_residual.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.ugx_tgrad.argtypes = [vp] + [dp] * 8
    L.ugx_tgrad_ooc.argtypes = [vp] + [dp] * 8
    pint = ctypes.POINTER(ctypes.c_int)
    L.ugx_pert_create.argtypes = [vp, dp, cd, pint]
    L.ugx_pert_solve.argtypes = [vp, ip, ip, ctypes.POINTER(Options), dp, pint, dp]
    L.ugx_pert_amplitudes.argtypes = [vp, ip, dp, dp, dp, dp]
    L.ugx_pert_residual.argtypes = [vp, ip, ip, dp, dp, dp, dp]
    L.ugx_linresp.argtypes = [vp, ip, ip, dp]
    L.ugx_pert_release.argtypes = [vp, ip]
    L.ugx_density.argtypes = [vp] + [dp] * 11
    L.ugx_set_profiling.argtypes = [vp, ip]
    L.ugx_set_profiling.restype = None
    L.ugx_kernel_stats.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    L.ugx_kernel_stats.restype = ctypes.c_size_t
    L.ugx_gemm.argtypes = [vp, ctypes.c_char, ctypes.c_char, ip, ip, ip, cd, vp, ctypes.c_long, vp, ctypes.c_long,
                           cd, vp, ctypes.c_long, ip, ip]
    L.ugx_gemm_bf16x3_static.argtypes = [vp, ctypes.c_char, ctypes.c_char, ip, ip, ip, cd, vp, ctypes.c_long, vp,
                                         ctypes.c_long, cd, vp, ctypes.c_long, ip]
    L.ugx_synchronize.argtypes = [vp]
    L.ugx_batch_solve.argtypes = [vp, ip, ctypes.POINTER(Problem), ctypes.POINTER(Options), ip,
                                  ctypes.POINTER(Result), dp]
    L.ugx_set_autotune.argtypes = [vp, ip]
    L.ugx_tile_table.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    L.ugx_tile_table.restype = ctypes.c_size_t
    L.ugx_save_tile_table.argtypes = [vp, ctypes.c_char_p]
    L.ugx_load_tile_table.argtypes = [vp, ctypes.c_char_p]
    fp, lp, cl = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_long), ctypes.c_long
    L.ugx_mlp_num_params.argtypes = []
    L.ugx_mlp_num_params.restype = cl
    L.ugx_mlp_create.argtypes = [ctypes.POINTER(vp), ip, ctypes.c_ulonglong]
    L.ugx_mlp_destroy.argtypes = [vp]
    L.ugx_mlp_destroy.restype = None
    L.ugx_mlp_last_error.argtypes = [vp]
    L.ugx_mlp_last_error.restype = ctypes.c_char_p
    L.ugx_mlp_set_params.argtypes = [vp, fp]
    L.ugx_mlp_get_params.argtypes = [vp, fp]
    L.ugx_mlp_set_scaling.argtypes = [vp, fp, fp, cd]
    L.ugx_mlp_forward.argtypes = [vp, fp, cl, ip, fp]
    L.ugx_mlp_loss_grad.argtypes = [vp, fp, fp, cl, dp, fp]
    L.ugx_mlp_train.argtypes = [vp, fp, fp, cl, ip, ip, cd, ctypes.c_ulonglong, dp]
    L.ugx_features.argtypes = [vp, cl, lp, fp]
    L.ugx_set_guess_network.argtypes = [vp, vp]
    non_int = ("ugx_default_options", "ugx_destroy", "ugx_last_error", "ugx_set_profiling", "ugx_kernel_stats",
               "ugx_tile_table",
               "ugx_mlp_num_params", "ugx_mlp_destroy", "ugx_mlp_last_error")
    for name in EXPORTS:
        if name not in non_int:
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def default_options():
    L = load_library()
    o = Options()
    L.ugx_default_options(ctypes.byref(o))
    return o


def split_blocks(ints, no):
    """Six Dirac blocks from a full nact^4 <pq|rs> array (occupied first)."""
    o, v = slice(0, no), slice(no, ints.shape[0])
    return dict(oooo=_c64(ints[o, o, o, o]), ooov=_c64(ints[o, o, o, v]), oovv=_c64(ints[o, o, v, v]),
                ovov=_c64(ints[o, v, o, v]), ovvv=_c64(ints[o, v, v, v]), vvvv=_c64(ints[v, v, v, v]))


class Context:
    """One ugx_ctx: a HIP stream on one device with a resident Hamiltonian."""

    def __init__(self, device=0):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.ugx_create(ctypes.byref(h), device)
        if rc != OK:
            raise UgxError(rc, "ugx_create failed with status %d (no usable GPU?)" % rc)
        self.h = h
        self.no = self.nv = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.ugx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, allow=(OK,)):
        if rc not in allow:
            raise UgxError(rc, self.L.ugx_last_error(self.h).decode())
        return rc

    def last_error(self):
        return self.L.ugx_last_error(self.h).decode()

    def load_hamiltonian(self, no, nv, fock, blocks):
        self._keep = [_c64(fock)] + [blocks[k] for k in ("oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")]
        self._check(self.L.ugx_load_hamiltonian(self.h, no, nv, *[_p(_c64(x)) for x in self._keep]))
        self._keep = None
        self.no, self.nv = no, nv

    def set_amplitudes(self, t1, t2):
        t1, t2 = _c64(t1), _c64(t2)
        self._check(self.L.ugx_set_amplitudes(self.h, _p(t1), _p(t2)))

    def amplitudes(self):
        t1 = np.zeros((self.no, self.nv))
        t2 = np.zeros((self.no, self.no, self.nv, self.nv))
        self._check(self.L.ugx_get_amplitudes(self.h, _p(t1), _p(t2)))
        return t1, t2

    def prepare(self, options=None, **kw):
        """ugx_prepare: allocations and conversions of a solve with these options."""
        o = options if options is not None else default_options()
        for k, v in kw.items():
            setattr(o, k, v)
        self._check(self.L.ugx_prepare(self.h, ctypes.byref(o)))

    def ccsd(self, options=None, **kw):
        o = options if options is not None else default_options()
        for k, v in kw.items():
            setattr(o, k, v)
        r = Result()
        rc = self._check(self.L.ugx_ccsd(self.h, ctypes.byref(o), ctypes.byref(r)), allow=(OK, NOT_CONVERGED))
        d = r.as_dict()
        d["converged"] = rc == OK
        return d

    def triples(self, precision=PREC_FP64):
        e = ctypes.c_double()
        self._check(self.L.ugx_triples(self.h, precision, ctypes.byref(e)))
        return e.value

    def triples_per_triplet(self, precision=PREC_FP64):
        n = self.no * (self.no + 1) * (self.no + 2) // 6
        e = np.zeros(n)
        self._check(self.L.ugx_triples_per_triplet(self.h, precision, _p(e)))
        return e

    def w3(self, i, j, k, precision=PREC_FP64):
        W = np.zeros((self.nv,) * 3)
        self._check(self.L.ugx_w3(self.h, i, j, k, precision, _p(W)))
        return W

    def energy(self, form=ENERGY_REFERENCE):
        e = ctypes.c_double()
        self._check(self.L.ugx_energy(self.h, form, ctypes.byref(e)))
        return e.value

    def residual(self, t1, t2, precision=PREC_FP64):
        t1, t2 = _c64(t1), _c64(t2)
        R1 = np.zeros_like(t1)
        R2 = np.zeros_like(t2)
        self._check(self.L.ugx_residual(self.h, _p(t1), _p(t2), precision, _p(R1), _p(R2)))
        return R1, R2

    def intermediates(self):
        o, v = self.no, self.nv
        out = dict(Fvv=np.zeros((v, v)), Foo=np.zeros((o, o)), Fov=np.zeros((o, v)),
                   Woooo=np.zeros((o, o, o, o)), Wovov=np.zeros((o, v, o, v)),
                   Wovvo=np.zeros((o, v, v, o)), Wovoo=np.zeros((o, v, o, o)))
        self._check(self.L.ugx_intermediates(self.h, *[_p(out[k]) for k in
                                                       ("Fvv", "Foo", "Fov", "Woooo", "Wovov", "Wovvo", "Wovoo")]))
        return out

    HBAR_BLOCKS = ("Hov", "Hoo", "Hvv", "Hoooo", "Hvvvv", "Hvovv", "Hooov", "Hovvo", "Hovov", "Hvvvo", "Hovoo")

    def hbar(self, which=None):
        """ugx_hbar: the HBAR blocks (hbar.cc:8-228) of the current amplitudes,
        as a dict of the blocks named in `which` (default all)."""
        o, v = self.no, self.nv
        shapes = dict(Hov=(o, v), Hoo=(o, o), Hvv=(v, v), Hoooo=(o, o, o, o), Hvvvv=(v, v, v, v),
                      Hvovv=(v, o, v, v), Hooov=(o, o, o, v), Hovvo=(o, v, v, o), Hovov=(o, v, o, v),
                      Hvvvo=(v, v, v, o), Hovoo=(o, v, o, o))
        which = self.HBAR_BLOCKS if which is None else which
        out = {k: np.zeros(shapes[k]) for k in which}
        self._check(self.L.ugx_hbar(self.h, *[_p(out[k]) if k in out else None for k in self.HBAR_BLOCKS]))
        return out

    def lambda_solve(self, options=None, **kw):
        """ugx_lambda: the CCSD Lambda equations (cclambda.cc:93-135) over the
        HBAR of the current amplitudes.  Returns converged, pseudoenergy,
        niter, rms and the pseudoenergy trace."""
        o = options if options is not None else default_options()
        for k, v in kw.items():
            setattr(o, k, v)
        pe, rms, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        trace = np.zeros(o.maxiter + 1)
        rc = self._check(self.L.ugx_lambda(self.h, ctypes.byref(o), ctypes.byref(pe), ctypes.byref(n),
                                           ctypes.byref(rms), _p(trace)), allow=(OK, NOT_CONVERGED))
        return dict(converged=rc == OK, pseudoenergy=pe.value, niter=n.value, rms=rms.value,
                    trace=trace[:n.value + 1])

    def lambdas(self):
        """ugx_get_lambda: (l1, l2) of the last lambda_solve."""
        l1 = np.zeros((self.no, self.nv))
        l2 = np.zeros((self.no, self.no, self.nv, self.nv))
        self._check(self.L.ugx_get_lambda(self.h, _p(l1), _p(l2)))
        return l1, l2

    def lambda_residual(self, l1, l2):
        """ugx_lambda_residual: (R1, R2, Goo, Gvv) at (l1, l2) over the HBAR
        of the current amplitudes (cclambda.cc:137-268)."""
        o, v = self.no, self.nv
        l1, l2 = _c64(l1), _c64(l2)
        R1, R2 = np.zeros((o, v)), np.zeros((o, o, v, v))
        Goo, Gvv = np.zeros((o, o)), np.zeros((v, v))
        self._check(self.L.ugx_lambda_residual(self.h, _p(l1), _p(l2), _p(R1), _p(R2), _p(Goo), _p(Gvv)))
        return R1, R2, Goo, Gvv

    TGRAD_BLOCKS = ("s1", "s2", "Doo", "Dvv", "Dov", "Gooov", "Gvvvo", "Goovv")

    def tgrad(self, which=None, ooc=False):
        """ugx_tgrad: the (T) gradient terms of CCSD(T) (ccwfn.cc:854-1098) of
        the current amplitudes: the Lambda sources s1, s2 and the (T) density
        parts named in `which` (default all).  A later lambda_solve with
        wfn_ccsd_t=1 adds the sources, and the density of that Lambda adds
        the (T) parts.  ooc=True: ugx_tgrad_ooc (ccwfn.cc:1150-1316, the
        OOC=true form: the (T) Doo and Dvv keep their diagonals only)."""
        o, v = self.no, self.nv
        shapes = dict(s1=(o, v), s2=(o, o, v, v), Doo=(o, o), Dvv=(v, v), Dov=(o, v), Gooov=(o, o, o, v),
                      Gvvvo=(v, v, v, o), Goovv=(o, o, v, v))
        which = self.TGRAD_BLOCKS if which is None else which
        out = {k: np.zeros(shapes[k]) for k in which}
        fn = self.L.ugx_tgrad_ooc if ooc else self.L.ugx_tgrad
        self._check(fn(self.h, *[_p(out[k]) if k in out else None for k in self.TGRAD_BLOCKS]))
        return out

    def pert(self, A, omega=0.0):
        """ugx_pert_create: a perturbation slot for the one-electron MO
        operator A (nact x nact) at omega (ccpert.cc:11-176), over the HBAR
        and Lambda amplitudes of the last lambda_solve."""
        return Perturbation(self, A, omega)

    DENSITY_BLOCKS = ("Doo", "Dvv", "Dov", "Dvo", "Goooo", "Gvvvv", "Gooov", "Gvvvo", "Govov", "Goovv")
    DENSITY_PARTS = ("oooo", "vvvv", "ooov", "vvvo", "ovov", "oovv")

    def density(self, which=None):
        """ugx_density: CCSD densities (ccdensity.cc:89-339) of the current
        amplitudes and the last Lambda solve (with the (T) parts when that
        solve included the (T) sources): e_one, e_two, the six parts of e_two
        and the blocks named in `which` (default all)."""
        o, v = self.no, self.nv
        shapes = dict(Doo=(o, o), Dvv=(v, v), Dov=(o, v), Dvo=(v, o), Goooo=(o, o, o, o), Gvvvv=(v, v, v, v),
                      Gooov=(o, o, o, v), Gvvvo=(v, v, v, o), Govov=(o, v, o, v), Goovv=(o, o, v, v))
        which = self.DENSITY_BLOCKS if which is None else which
        blocks = {k: np.zeros(shapes[k]) for k in which}
        E = np.zeros(8)
        self._check(self.L.ugx_density(self.h, _p(E), *[_p(blocks[k]) if k in blocks else None
                                                        for k in self.DENSITY_BLOCKS]))
        return dict(e_one=E[0], e_two=E[1], parts=dict(zip(self.DENSITY_PARTS, E[2:])), blocks=blocks)

    def set_profiling(self, on=True):
        self.L.ugx_set_profiling(self.h, 1 if on else 0)

    def kernel_stats(self):
        n = self.L.ugx_kernel_stats(self.h, None, 0)
        buf = ctypes.create_string_buffer(max(n, 1))
        self.L.ugx_kernel_stats(self.h, buf, n)
        out = {}
        for line in buf.value.decode().splitlines():
            name, launches, ms, gflop = line.rsplit(" ", 3)
            out[name] = dict(launches=int(launches), ms=float(ms), gflop=float(gflop))
        return out

    def gemm(self, opa, opb, M, N, K, alpha, A_ptr, lda, B_ptr, ldb, beta, C_ptr, ldc, precision=PREC_FP64,
             tile=-1):
        self._check(self.L.ugx_gemm(self.h, opa.encode(), opb.encode(), M, N, K, alpha, A_ptr, lda, B_ptr, ldb,
                                    beta, C_ptr, ldc, precision, tile))

    def gemm_bf16x3_static(self, opa, opb, M, N, K, alpha, A_ptr, lda, B_ptr, ldb, beta, C_ptr, ldc, tile=-1):
        """ugx_gemm_bf16x3_static: BF16x3 GEMM of FP32 device operands with the
        library's static staging array (no device allocation)."""
        self._check(self.L.ugx_gemm_bf16x3_static(self.h, opa.encode(), opb.encode(), M, N, K, alpha, A_ptr, lda,
                                                  B_ptr, ldb, beta, C_ptr, ldc, tile))

    def synchronize(self):
        self._check(self.L.ugx_synchronize(self.h))

    def batch_solve(self, problems, options=None, nworkers=8, **kw):
        """Solves [(no, nv, fock, blocks), ...] with ugx_batch_solve; returns
        (list of result dicts, batch wall seconds)."""
        o = options if options is not None else default_options()
        for k, v in kw.items():
            setattr(o, k, v)
        keep = []
        arr = (Problem * len(problems))()
        for q, (no, nv, fock, blocks) in enumerate(problems):
            bufs = [_c64(fock)] + [_c64(blocks[k]) for k in ("oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")]
            keep.append(bufs)
            arr[q].no, arr[q].nv = no, nv
            for name, b in zip(("fock", "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv"), bufs):
                setattr(arr[q], name, _p(b))
        res = (Result * len(problems))()
        wall = ctypes.c_double()
        rc = self._check(self.L.ugx_batch_solve(self.h, len(problems), arr, ctypes.byref(o), nworkers, res,
                                                ctypes.byref(wall)), allow=(OK, NOT_CONVERGED))
        out = []
        for r in res:
            d = r.as_dict()
            d["converged"] = rc == OK
            out.append(d)
        return out, wall.value

    def set_autotune(self, on=True):
        self._check(self.L.ugx_set_autotune(self.h, 1 if on else 0))

    def tile_table(self):
        n = self.L.ugx_tile_table(self.h, None, 0)
        buf = ctypes.create_string_buffer(max(n, 1))
        self.L.ugx_tile_table(self.h, buf, n)
        return buf.value.decode()

    def save_tile_table(self, path):
        self._check(self.L.ugx_save_tile_table(self.h, os.fsencode(path)))

    def load_tile_table(self, path):
        self._check(self.L.ugx_load_tile_table(self.h, os.fsencode(path)))

    def features(self, idx=None, n=None):
        """Raw learned-guess features of the amplitudes idx (flattened ijab)."""
        if idx is not None:
            idx = np.ascontiguousarray(idx, dtype=np.int64)
            n = idx.size
        elif n is None:
            n = self.no * self.no * self.nv * self.nv
        X = np.zeros((n, MLP_NFEAT), dtype=np.float32)
        ip = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_long)) if idx is not None else None
        self._check(self.L.ugx_features(self.h, n, ip, _f(X)))
        return X

    def set_guess_network(self, net):
        """Network used by GUESS_MLP (None clears it); the context keeps a reference."""
        self._net = net
        self._check(self.L.ugx_set_guess_network(self.h, net.h if net is not None else None))


class Perturbation:
    """A perturbation slot of a context (ugx_pert_*): the perturbed wave
    function of ccpert.cc and its linear response (ccresp.cc)."""

    def __init__(self, ctx, A, omega):
        self.ctx = ctx
        self.L = ctx.L
        n = ctx.no + ctx.nv
        A = _c64(A)
        if A.shape != (n, n):
            raise ValueError("perturbation must be nact x nact")
        h = ctypes.c_int()
        ctx._check(self.L.ugx_pert_create(ctx.h, _p(A), float(omega), ctypes.byref(h)))
        self.h = h.value
        self.omega = omega

    def solve(self, hand=0, options=None, **kw):
        """ugx_pert_solve: hand 0 right (X), 1 left (Y)."""
        o = options if options is not None else default_options()
        for k, v in kw.items():
            setattr(o, k, v)
        pr, rms, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        rc = self.ctx._check(self.L.ugx_pert_solve(self.ctx.h, self.h, hand, ctypes.byref(o), ctypes.byref(pr),
                                                   ctypes.byref(n), ctypes.byref(rms)), allow=(OK, NOT_CONVERGED))
        return dict(converged=rc == OK, pseudoresponse=pr.value, niter=n.value, rms=rms.value)

    def amplitudes(self):
        no, nv = self.ctx.no, self.ctx.nv
        X1, Y1 = np.zeros((no, nv)), np.zeros((no, nv))
        X2, Y2 = np.zeros((no, no, nv, nv)), np.zeros((no, no, nv, nv))
        self.ctx._check(self.L.ugx_pert_amplitudes(self.ctx.h, self.h, _p(X1), _p(X2), _p(Y1), _p(Y2)))
        return X1, X2, Y1, Y2

    def residual(self, hand, A1, A2):
        no, nv = self.ctx.no, self.ctx.nv
        A1, A2 = _c64(A1), _c64(A2)
        R1, R2 = np.zeros((no, nv)), np.zeros((no, no, nv, nv))
        self.ctx._check(self.L.ugx_pert_residual(self.ctx.h, self.h, hand, _p(A1), _p(A2), _p(R1), _p(R2)))
        return R1, R2

    def linresp(self, other):
        """ugx_linresp: <<A;B>> with A this perturbation, B other's."""
        v = ctypes.c_double()
        self.ctx._check(self.L.ugx_linresp(self.ctx.h, self.h, other.h, ctypes.byref(v)))
        return v.value

    def release(self):
        """ugx_pert_release.  The handle is kept: later calls with it are
        rejected by the library (UGX_BAD_INPUT)."""
        self.ctx._check(self.L.ugx_pert_release(self.ctx.h, self.h))


class Network:
    """One ugx_mlp: the learned T2 guess network (parameters, Adam state) on a device."""

    def __init__(self, device=0, seed=0):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.ugx_mlp_create(ctypes.byref(h), device, seed)
        if rc != OK:
            raise UgxError(rc, "ugx_mlp_create failed with status %d" % rc)
        self.h = h
        self.nparam = int(self.L.ugx_mlp_num_params())
        self.shift = np.zeros(MLP_NFEAT, np.float32)
        self.scale = np.ones(MLP_NFEAT, np.float32)
        self.y_scale = 1.0

    def close(self):
        if getattr(self, "h", None):
            self.L.ugx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise UgxError(rc, self.L.ugx_mlp_last_error(self.h).decode())

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        self._check(self.L.ugx_mlp_get_params(self.h, _f(p)))
        return p

    def set_params(self, p):
        p = _c32(p)
        assert p.size == self.nparam
        self._check(self.L.ugx_mlp_set_params(self.h, _f(p)))

    def set_scaling(self, shift, scale, y_scale):
        self.shift, self.scale, self.y_scale = _c32(shift), _c32(scale), float(y_scale)
        self._check(self.L.ugx_mlp_set_scaling(self.h, _f(self.shift), _f(self.scale), self.y_scale))

    def forward(self, X, precision=PREC_FP32):
        X = _c32(X)
        y = np.zeros(X.shape[0], np.float32)
        self._check(self.L.ugx_mlp_forward(self.h, _f(X), X.shape[0], precision, _f(y)))
        return y

    def loss_grad(self, X, y):
        X, y = _c32(X), _c32(y)
        loss = ctypes.c_double()
        g = np.zeros(self.nparam, np.float32)
        self._check(self.L.ugx_mlp_loss_grad(self.h, _f(X), _f(y), X.shape[0], ctypes.byref(loss), _f(g)))
        return loss.value, g

    def train(self, X, y, epochs, batch, lr, seed=0):
        X, y = _c32(X), _c32(y)
        hist = np.zeros(epochs)
        self._check(self.L.ugx_mlp_train(self.h, _f(X), _f(y), X.shape[0], epochs, batch, lr, seed, _p(hist)))
        return hist

    def save(self, path):
        np.savez(path, params=self.params(), shift=self.shift, scale=self.scale, y_scale=np.float64(self.y_scale))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_scaling(z["shift"], z["scale"], float(z["y_scale"]))
            self.set_params(z["params"])
        return self
