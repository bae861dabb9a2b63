"""The dump / fixture format (ugacc_amd/dump.py, modelled on the items of
attic/ccdump.cc:27-67) and the committed golden fixtures.

- write -> read round trip of sizes, scalars, generator and arrays (bitwise);
- corrupted bytes, a bad magic and a wrong rebuilt array are detected;
- every committed fixture reads with its checksums intact;
- the synthetic fixtures' generators rebuild their stored Hamiltonians
  bitwise, and the oracle reproduces their stored results (guarding both
  the fixtures and the oracle against drift).
"""
import glob
import os

import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd import dump
from ugacc_amd.hamiltonian import blocks_from_full, full_from_blocks
from ugacc_amd.synthetic import random_hamiltonian

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "*.ugxd")))
SYN = [p for p in FIXTURES if os.path.basename(p).startswith("syn_")]


def test_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    arrays = dict(t1=rng.standard_normal((3, 7)), t2=rng.standard_normal((3, 3, 7, 7)), e=np.array(1.5),
                  empty=np.zeros((0, 4)))
    big = rng.standard_normal((5, 6))
    p = str(tmp_path / "x.ugxd")
    dump.write(p, dict(no=3, nv=7, nfzc=1), dict(escf=-76.5, eccsd=-0.2), arrays, dict(kind="test", seed=9),
               described=dict(vvvv=big))
    d = dump.Dump(p)
    assert d.sizes == dict(no=3, nv=7, nfzc=1)
    assert d.scalars == dict(escf=-76.5, eccsd=-0.2)
    assert d.generator == dict(kind="test", seed=9)
    assert sorted(d.arrays) == sorted(arrays)
    for k, a in arrays.items():
        b = d[k]
        assert b.shape == a.shape and np.array_equal(a, b)
    assert d.verify("vvvv", big)
    assert not d.verify("vvvv", np.nextafter(big, np.inf))  # one ulp anywhere changes the digest
    assert d.verify("vvvv", big * (1 + 1e-15), rtol=1e-12)
    with pytest.raises(dump.DumpError):
        d.verify("vvvv", big.T)
    # the data region starts 8-aligned and every array offset is a multiple of 8
    raw = open(p, "rb").read()
    hlen = int.from_bytes(raw[8:16], "little")
    assert hlen % 8 == 0
    assert all(r["offset"] % 8 == 0 for r in d.header["arrays"].values())


def test_corruption_is_detected(tmp_path):
    p = str(tmp_path / "x.ugxd")
    a = np.arange(64, dtype=float).reshape(8, 8)
    dump.write(p, dict(no=1, nv=1), {}, dict(a=a))
    raw = bytearray(open(p, "rb").read())
    raw[-3] ^= 0x01  # flip one bit of the last element
    open(p, "wb").write(bytes(raw))
    with pytest.raises(dump.DumpError, match="checksum"):
        dump.Dump(p)["a"]
    raw[:8] = b"NOTADUMP"
    open(p, "wb").write(bytes(raw))
    with pytest.raises(dump.DumpError, match="magic"):
        dump.Dump(p)


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_committed_fixtures_read_with_checksums(path):
    d = dump.Dump(path)
    assert d.generator.get("kind") in ("synthetic", "molecule")
    assert "seed" in d.generator or d.generator.get("builder", "").endswith("water")
    for name in d.arrays:
        assert np.all(np.isfinite(d[name]))  # every SHA-256 checked on read
    no, nv = d.sizes["no"], d.sizes["nv"]
    assert d["t1"].shape == (no, nv)
    assert d.scalars["eccsd"] < 0 and d.scalars["e_mp2"] < 0


@pytest.mark.parametrize("path", SYN, ids=os.path.basename)
def test_synthetic_fixtures_rebuild_and_oracle_reproduces(path):
    d = dump.Dump(path)
    g = dict(d.generator)
    no, nv = g.pop("no"), g.pop("nv")
    seed = g.pop("seed")
    g.pop("kind"), g.pop("module")
    fock, ints = random_hamiltonian(no, nv, seed=seed, **g)
    f_d, blocks_d = d.hamiltonian()
    assert np.array_equal(fock, f_d)
    for k, b in blocks_from_full(ints, no).items():
        assert np.array_equal(b, blocks_d[k]), k
    # the generator's 8-fold symmetry holds to rounding (a BLAS Gram product)
    np.testing.assert_allclose(full_from_blocks(no, nv, blocks_d), ints, rtol=0, atol=1e-16)
    o = Oracle(f_d, ints, no)
    t1g, t2g = o.amps()
    assert np.array_equal(t1g, d["t1_guess"]) and np.array_equal(t2g, d["t2_guess"])
    assert o.energy() == d.scalars["e_mp2"]
    _, R1, R2 = o.iterate()
    np.testing.assert_allclose(R1, d["R1_iter1"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(R2, d["R2_iter1"], rtol=0, atol=1e-15)
    for k, v in o.intermediates().items():
        np.testing.assert_allclose(v, d[k + "_iter1"], rtol=0, atol=1e-15, err_msg=k)
    o2 = Oracle(f_d, ints, no)
    r = o2.ccsd(maxiter=200, conv=d.scalars["r_convergence"], diis=True, mode=0, diis_policy=0, min_vecs=3)
    assert r["niter"] == d.scalars["niter"]
    assert abs(r["eccsd"] - d.scalars["eccsd"]) < 1e-13
    t1, t2 = o2.amps()
    assert np.max(np.abs(t1 - d["t1"])) < 1e-13 and np.max(np.abs(t2 - d["t2"])) < 1e-13
    # the reference's (T) values: tcorr = tcorr_ooc = the folded per-triplet sum
    assert abs(d.scalars["e_t_tcorr"] - d.scalars["e_t_tcorr_ooc"]) < 1e-14
    assert abs(d["e_triplet"].sum() - d.scalars["e_t_tcorr"]) < 1e-14
