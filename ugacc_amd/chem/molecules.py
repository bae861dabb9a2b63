"""Geometries of the benchmark systems (bohr).

- H2O at O-H 0.9572 A, H-O-H 104.52 deg (the water configuration);
- seeded water clusters: O atoms on a random cyclic ring with O-O distances
  drawn from U[2.70, 2.90] A, each monomer at the H2O geometry with a random
  rotation; draws with an intermolecular H...H < 1.5 A or a non-neighbour
  O...O < 2.6 A are rejected and redrawn;
- seeded distorted monomers (O-H in U[0.90, 1.10] A, angle in U[95, 115]
  deg) and dimers (O-O in U[2.7, 3.2] A) for the batched workload;
- the H2O and HOF geometries of the reference's saved runs
  (psi3/output.c1_fzc:28-31 and input.hof:9-11).
"""
import numpy as np

ANGSTROM = 1.0 / 0.52917721067  # bohr per angstrom (CODATA 2014, as Psi4)

PSI3_H2O = [("O", (0.0, 0.0, -0.116529192243)), ("H", (0.0, 1.344767972762, 0.924701422022)),
            ("H", (0.0, -1.344767972762, 0.924701422022))]
HOF = [("O", (0.0, 0.0, -0.116529192243)), ("H", (0.0, 1.344767972762, 0.924701422022)),
       ("F", (0.0, -1.344767972762, 0.924701422022))]


def water(r_oh=0.9572, angle=104.52):
    """Monomer in its own frame (angstrom): O at the origin, molecule in xz."""
    h = np.radians(angle) / 2
    return [("O", np.zeros(3)), ("H", r_oh * np.array([np.sin(h), 0.0, np.cos(h)])),
            ("H", r_oh * np.array([-np.sin(h), 0.0, np.cos(h)]))]


def to_bohr(atoms_ang):
    return [(s, tuple(float(x) * ANGSTROM for x in xyz)) for s, xyz in atoms_ang]


def random_rotation(rng):
    """Uniform SO(3) rotation from a random unit quaternion."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def _ring(dists):
    """Points of a convex cyclic polygon with the given side lengths."""
    dists = np.asarray(dists, dtype=float)
    lo, hi = dists.max() / 2 * (1 + 1e-12), dists.sum()
    f = lambda R: np.sum(2 * np.arcsin(np.clip(dists / (2 * R), -1, 1))) - 2 * np.pi
    for _ in range(200):  # bisection on the circumradius
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            lo = mid
        else:
            hi = mid
    R = 0.5 * (lo + hi)
    ang = np.concatenate([[0.0], np.cumsum(2 * np.arcsin(dists / (2 * R)))[:-1]])
    return np.stack([R * np.cos(ang), R * np.sin(ang), np.zeros_like(ang)], axis=1)


def water_cluster(n, seed, oo=(2.70, 2.90), max_tries=1000):
    """Seeded (H2O)n ring cluster, atoms in bohr."""
    rng = np.random.default_rng(seed)
    mono = water()
    for _ in range(max_tries):
        d = rng.uniform(oo[0], oo[1], n)
        O = _ring(d) if n > 2 else np.array([[0.0, 0.0, 0.0], [d[0], 0.0, 0.0]])
        atoms = []
        for k in range(n):
            Rm = random_rotation(rng)
            for s, xyz in mono:
                atoms.append((s, O[k] + Rm @ xyz))
        if _acceptable(atoms, n):
            return to_bohr(atoms)
    raise RuntimeError("no acceptable cluster geometry")


def _acceptable(atoms, n):
    pos = np.array([x for _, x in atoms])
    mol = np.repeat(np.arange(n), 3)
    sym = [s for s, _ in atoms]
    for i in range(len(atoms)):
        for j in range(i):
            if mol[i] == mol[j]:
                continue
            r = np.linalg.norm(pos[i] - pos[j])
            if sym[i] == "H" and sym[j] == "H" and r < 1.5:
                return False
            if sym[i] == "O" and sym[j] == "O" and n > 3 and abs(mol[i] - mol[j]) not in (1, n - 1) and r < 2.6:
                return False
    return True


def distorted_monomer(seed):
    rng = np.random.default_rng(seed)
    r1, r2 = rng.uniform(0.90, 1.10, 2)
    ang = np.radians(rng.uniform(95.0, 115.0))
    h = ang / 2
    atoms = [("O", np.zeros(3)), ("H", r1 * np.array([np.sin(h), 0.0, np.cos(h)])),
             ("H", r2 * np.array([-np.sin(h), 0.0, np.cos(h)]))]
    return to_bohr(atoms)


def random_dimer(seed):
    rng = np.random.default_rng(seed)
    mono = water()
    for _ in range(1000):
        d = rng.uniform(2.7, 3.2)
        atoms = []
        for k, O in enumerate([np.zeros(3), np.array([d, 0.0, 0.0])]):
            Rm = random_rotation(rng)
            for s, xyz in mono:
                atoms.append((s, O + Rm @ xyz))
        if _acceptable(atoms, 2):
            return to_bohr(atoms)
    raise RuntimeError("no acceptable dimer")
