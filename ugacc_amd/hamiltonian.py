"""Host-side Hamiltonian layouts (the counterpart of hamiltonian.cc:19-103).

The library keeps the six Dirac blocks oooo, ooov, oovv, ovov, ovvv, vvvv
(occupied first); the reference (and the oracle) keep the full nact^4 array
ints[p][q][r][s] = <pq|rs>.  Real orbitals give the 8-fold symmetry
<pq|rs> = (pr|qs), so every other o/v pattern of the full array is a
transpose of one of the six blocks.
"""
import itertools

import numpy as np

BLOCKS = ("oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")

# Position permutations pi with <x0 x1|x2 x3> = <x_pi0 x_pi1|x_pi2 x_pi3>,
# from the chemist symmetry (pr|qs) = (rp|qs) = (pr|sq) = (qs|pr) ...
SYMMETRIES = ((0, 1, 2, 3), (2, 1, 0, 3), (0, 3, 2, 1), (2, 3, 0, 1),
              (1, 0, 3, 2), (3, 0, 1, 2), (1, 2, 3, 0), (3, 2, 1, 0))


def blocks_from_full(ints, no):
    o, v = slice(0, no), slice(no, ints.shape[0])
    sl = {"o": o, "v": v}
    return {name: np.ascontiguousarray(ints[tuple(sl[c] for c in name)]) for name in BLOCKS}


def full_from_blocks(no, nv, blocks, out=None):
    """nact^4 array <pq|rs> assembled from the six blocks."""
    n = no + nv
    if out is None:
        out = np.empty((n, n, n, n))
    sl = {"o": slice(0, no), "v": slice(no, n)}
    for pattern in itertools.product("ov", repeat=4):
        for pi in SYMMETRIES:
            name = "".join(pattern[k] for k in pi)
            if name in blocks:
                # out[x0,x1,x2,x3] = B[x_pi0, x_pi1, x_pi2, x_pi3]
                inv = np.argsort(pi)
                out[tuple(sl[c] for c in pattern)] = blocks[name].transpose(inv)
                break
        else:
            raise ValueError("no canonical block for pattern " + "".join(pattern))
    return out
