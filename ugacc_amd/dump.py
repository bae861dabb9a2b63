"""Dump / fixture format of a CC problem and its results.

The reference's dump (attic/ccdump.cc:12-70) writes, to a PSIO file, the
sizes nmo, nact, no, nv, nfzc, nfzv, the energies escf, efzc, eccsd, and the
arrays T1, T2, the Fock matrix, the two-electron integrals <pq|rs> and L.
This format holds the same items in a self-describing, portable file:

    bytes 0..7      magic b"UGXDUMP1"
    bytes 8..15     H, the header length, little-endian uint64
    bytes 16..16+H  header: UTF-8 JSON, padded with spaces to a multiple of 8
    then            the arrays, each row-major little-endian float64,
                    starting at its header "offset" (from the end of the
                    header; every offset a multiple of 8)

Header keys:
    format, version           "ugacc-dump", 1
    sizes                     nmo, nact, no, nv, nfzc, nfzv (ccdump.cc:27-32)
    scalars                   energies and other numbers (escf, efzc, eccsd
                              as in ccdump.cc:34-36, plus e_mp2, e_t, ...)
    generator                 how the problem was made (e.g. the seeded
                              cluster generator and basis, or the synthetic
                              Hamiltonian's seed and scales), so that a
                              problem whose integrals are not stored can be
                              rebuilt
    arrays                    name -> {shape, offset, sha256} of every stored
                              array
    checksums                 name -> {shape, sha256, sum, norm} of arrays
                              that are described but not stored (large
                              integral blocks rebuilt from the generator)

Array names: "t1" [no][nv], "t2" [no][no][nv][nv], "fock" [nact][nact], the
six Dirac blocks "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv" in the
layouts of include/ugx.h (the integrals the reference stores as one
nact^4 array; L = 2<pq|rs> - <pq|sr> is derived from them,
hamiltonian.cc:96-102, and not stored), and any further named results
(intermediates, residuals, per-triplet energies).
"""
import hashlib
import json
import os
import struct

import numpy as np

MAGIC = b"UGXDUMP1"
VERSION = 1
BLOCKS = ("oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")


class DumpError(ValueError):
    pass


def _le64(a):
    return np.ascontiguousarray(a, dtype="<f8")


def digest(a):
    """SHA-256 of an array's little-endian FP64 bytes (row-major)."""
    return hashlib.sha256(_le64(a).tobytes()).hexdigest()


def summary(a):
    """Checksum record of an array that is described but not stored."""
    a = _le64(a)
    return dict(shape=list(a.shape), sha256=digest(a), sum=float(a.sum()), norm=float(np.sqrt((a * a).sum())))


def write(path, sizes, scalars=None, arrays=None, generator=None, described=None):
    """Writes a dump.  arrays: name -> ndarray (stored); described: name ->
    ndarray (only its checksum record is stored)."""
    arrays = arrays or {}
    for k in ("no", "nv"):
        if k not in sizes:
            raise DumpError("sizes needs %r" % k)
    hdr = dict(format="ugacc-dump", version=VERSION, sizes={k: int(v) for k, v in sizes.items()},
               scalars={k: float(v) for k, v in (scalars or {}).items()}, generator=generator or {},
               arrays={}, checksums={k: summary(v) for k, v in (described or {}).items()})
    data, off = [], 0
    for name, a in arrays.items():
        b = _le64(a).tobytes()
        hdr["arrays"][name] = dict(shape=list(np.shape(a)), offset=off, sha256=hashlib.sha256(b).hexdigest())
        data.append(b)
        off += len(b)  # float64 data keeps every offset a multiple of 8
    h = json.dumps(hdr, indent=1, sort_keys=True).encode()
    h += b" " * (-len(h) % 8)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<Q", len(h)))
        f.write(h)
        for b in data:
            f.write(b)
    os.replace(tmp, path)


class Dump:
    """A read dump: .sizes, .scalars, .generator, .checksums and the arrays
    (by name, .arrays lists the stored ones).  Every array read is checked
    against its SHA-256."""

    def __init__(self, path):
        with open(path, "rb") as f:
            raw = f.read()
        if raw[:8] != MAGIC:
            raise DumpError("%s: not a ugacc dump (bad magic)" % path)
        (hlen,) = struct.unpack("<Q", raw[8:16])
        if hlen % 8 or 16 + hlen > len(raw):
            raise DumpError("%s: bad header length" % path)
        hdr = json.loads(raw[16:16 + hlen].decode())
        if hdr.get("format") != "ugacc-dump" or hdr.get("version") != VERSION:
            raise DumpError("%s: unsupported format/version" % path)
        self.path = path
        self.header = hdr
        self.sizes = hdr["sizes"]
        self.scalars = hdr["scalars"]
        self.generator = hdr["generator"]
        self.checksums = hdr.get("checksums", {})
        self._data = memoryview(raw)[16 + hlen:]
        self.arrays = list(hdr["arrays"].keys())

    def __contains__(self, name):
        return name in self.header["arrays"]

    def __getitem__(self, name):
        rec = self.header["arrays"].get(name)
        if rec is None:
            raise KeyError(name)
        n = int(np.prod(rec["shape"])) if rec["shape"] else 1
        off = rec["offset"]
        if off % 8 or off + 8 * n > len(self._data):
            raise DumpError("%s: array %s out of bounds" % (self.path, name))
        b = bytes(self._data[off:off + 8 * n])
        if hashlib.sha256(b).hexdigest() != rec["sha256"]:
            raise DumpError("%s: checksum mismatch in %s" % (self.path, name))
        return np.frombuffer(b, dtype="<f8").reshape(rec["shape"]).astype(np.float64)

    def verify(self, name, a, rtol=0.0):
        """Checks an array rebuilt from the generator against the checksum
        record of `name`: bitwise (rtol 0) or its sum and norm to rtol."""
        rec = self.checksums[name]
        a = _le64(a)
        if list(a.shape) != rec["shape"]:
            raise DumpError("%s: shape %s, recorded %s" % (name, list(a.shape), rec["shape"]))
        if rtol == 0.0:
            return digest(a) == rec["sha256"]
        s = summary(a)
        return (abs(s["norm"] - rec["norm"]) <= rtol * max(1.0, rec["norm"])
                and abs(s["sum"] - rec["sum"]) <= rtol * max(1.0, rec["norm"]))

    def hamiltonian(self):
        """(fock, blocks) when the dump stores them."""
        return self["fock"], {k: self[k] for k in BLOCKS}
