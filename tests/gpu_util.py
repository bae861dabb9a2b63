"""Shared helpers for the GPU tests (marked ``gpu``): one library context per
session and seeded synthetic Hamiltonians."""
import numpy as np
import pytest

from ugacc_amd import ugx
from ugacc_amd.synthetic import random_hamiltonian


_ctx = None


def context():
    global _ctx
    if _ctx is None:
        _ctx = ugx.Context(0)
    return _ctx


def loaded(no, nv, seed, **kw):
    """(ctx, fock, ints) with the synthetic Hamiltonian loaded on the device."""
    fock, ints = random_hamiltonian(no, nv, seed=seed, **kw)
    c = context()
    c.load_hamiltonian(no, nv, fock, ugx.split_blocks(ints, no))
    return c, fock, ints


def announce(msg):
    print("[gpu-case] " + msg, flush=True)
