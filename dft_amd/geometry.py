"""Synthetic geometries of the workloads: the benzene ring and seeded random
molecules and clusters (nuclei as [Z, x, y, z] rows, bohr).

The reference builds its molecules by hand in main.cpp (Molecule::addNucleus,
Molecule.cpp:10-56); these generators only produce such lists.  Every
generator is a pure function of its arguments (numpy's PCG64 stream seeded
with the given seed), so a seed names one geometry on every machine."""
import numpy as np

TRAINING_ELEMENTS = (1, 6, 7, 8)


def benzene():
    """Planar ring in z = 0: C at radius 2.64 bohr (C-C = 2.64), H at 2.64 + 2.05 = 4.69."""
    out = []
    for k in range(6):
        a = np.pi / 3 * k
        out.append([6, 2.64 * np.cos(a), 2.64 * np.sin(a), 0.0])
    for k in range(6):
        a = np.pi / 3 * k
        out.append([1, 4.69 * np.cos(a), 4.69 * np.sin(a), 0.0])
    return out


def random_cluster(seed, natoms, radius, zs=TRAINING_ELEMENTS, min_dist=2.0, ne_range=None):
    """Seeded random molecule: Z uniform in zs, positions uniform in a ball,
    pairwise distance >= min_dist; an H is added when the electron count is
    odd; ne_range rejects clusters outside [lo, hi] electrons."""
    rng = np.random.default_rng(seed)
    while True:
        pos = []
        while len(pos) < natoms:
            p = rng.uniform(-radius, radius, 3)
            if np.linalg.norm(p) > radius:
                continue
            if all(np.linalg.norm(p - q) >= min_dist for q in pos):
                pos.append(p)
        Z = rng.choice(zs, natoms)
        atoms = [[float(z), *p] for z, p in zip(Z, pos)]
        ne = int(sum(Z))
        if ne % 2:
            for _ in range(1000):
                p = rng.uniform(-radius, radius, 3)
                if np.linalg.norm(p) <= radius and all(np.linalg.norm(p - np.array(a[1:])) >= min_dist for a in atoms):
                    atoms.append([1.0, *p])
                    ne += 1
                    break
        if ne % 2 == 0 and (ne_range is None or ne_range[0] <= ne <= ne_range[1]):
            return atoms


def training_molecule(seed, min_atoms=4, max_atoms=12, radius=5.0):
    """One molecule of the learned preconditioner's set: min_atoms..max_atoms
    atoms (count drawn from the seed), Z uniform in {H, C, N, O}, uniform in a
    ball of the given radius (a 10-bohr sphere read as its diameter), pair
    distances >= 2 bohr, an even electron count (an H added when odd)."""
    natoms = int(np.random.default_rng([seed, 1]).integers(min_atoms, max_atoms + 1))
    return random_cluster(seed, natoms, radius)
