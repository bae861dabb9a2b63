"""ctypes binding of integrals.cpp (built in-tree with g++ -fopenmp)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libugxchem.so")
SRC = os.path.join(HERE, "integrals.cpp")

_lib = None


def build():
    cmd = ["g++", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-o", LIB + ".tmp", SRC]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = ctypes.CDLL(LIB)
        dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.chem_nbf.argtypes = [ctypes.c_int, ip]
        L.chem_one_electron.argtypes = [ctypes.c_int, dp, ip, ip, ip, dp, dp, ctypes.c_int, dp, dp, dp, dp, dp]
        L.chem_eri.argtypes = [ctypes.c_int, dp, ip, ip, ip, dp, dp, ctypes.c_double, dp]
        L.chem_properties.argtypes = [ctypes.c_int, dp, ip, ip, ip, dp, dp, dp, dp, dp]
        L.chem_jk.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.chem_jk.restype = None
        L.chem_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class ShellArrays:
    def __init__(self, shells):
        self.nsh = len(shells)
        self.xyz = np.ascontiguousarray([c for c, _, _ in shells], dtype=np.float64).ravel()
        self.l = np.array([l for _, l, _ in shells], dtype=np.int32)
        self.nprim = np.array([len(p) for _, _, p in shells], dtype=np.int32)
        self.off = np.concatenate([[0], np.cumsum(self.nprim)[:-1]]).astype(np.int32)
        self.exps = np.array([e for _, _, p in shells for e, _ in p], dtype=np.float64)
        self.coefs = np.array([c for _, _, p in shells for _, c in p], dtype=np.float64)
        self.nbf = int(lib().chem_nbf(self.nsh, _ip(self.l)))

    def args(self):
        return (self.nsh, _dp(self.xyz), _ip(self.l), _ip(self.nprim), _ip(self.off), _dp(self.exps),
                _dp(self.coefs))


def one_electron(shells, charges, coords):
    sa = ShellArrays(shells)
    n = sa.nbf
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    Z = np.ascontiguousarray(charges, dtype=np.float64)
    C = np.ascontiguousarray(coords, dtype=np.float64).ravel()
    rc = lib().chem_one_electron(*sa.args(), len(Z), _dp(Z), _dp(C), _dp(S), _dp(T), _dp(V))
    if rc != 0:
        raise ValueError("angular momentum above d")
    return S, T, V


def eri(shells, thresh=1e-14):
    sa = ShellArrays(shells)
    n = sa.nbf
    out = np.zeros((n, n, n, n))
    rc = lib().chem_eri(*sa.args(), thresh, _dp(out))
    if rc != 0:
        raise ValueError("angular momentum above d")
    return out


def jk(eri_ao, D):
    n = D.shape[0]
    D = np.ascontiguousarray(D)
    J, K = np.zeros((n, n)), np.zeros((n, n))
    lib().chem_jk(n, _dp(eri_ao), _dp(D), _dp(J), _dp(K))
    return J, K


def properties(shells, origin=(0.0, 0.0, 0.0)):
    """AO dipole integrals <p|(r - O)|q> and angular-momentum integrals
    <p|(r - O) x grad|q> (real, antisymmetric), each (3, nbf, nbf)."""
    sa = ShellArrays(shells)
    n = sa.nbf
    D, M = np.zeros((3, n, n)), np.zeros((3, n, n))
    O = np.ascontiguousarray(origin, dtype=np.float64)
    rc = lib().chem_properties(*sa.args(), _dp(O), _dp(D), _dp(M))
    if rc != 0:
        raise ValueError("angular momentum above d")
    return D, M
