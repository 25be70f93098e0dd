"""Write the C2 SCF-step fixture (tests/golden/c2_step.npz) from the compiled reference.

The neon workload's grid (96^3, h = 0.2, bounds +-9.5), PBE96, closed shell (5 alpha + 5 beta
electrons): one executeMolecularSCFIteration (MolecularSCFIteration.cpp:12-252) of the reference,
reference Poisson rule and seed-vector Davidson starts, on the input spin densities
c2_input_density() below.  Stored: the energy components, the eigenvalues, dRho, the KS residual,
and every field (V_H, V_xc, the output densities, the output V_H) at 4096 seeded grid points plus
its sum and sum of squares.  Slow (minutes of CPU: the reference's boundary pass alone is O(n^5));
run it where the reference checkout exists, after `make -C oracle ref`.
Usage: python tests/golden/make_c2_step.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import Grid, Reference  # noqa: E402

NUCLEI = [[10.0, 0.0, 0.0, 0.0]]
FUNCTIONAL = 2


def c2_grid():
    return Grid.centered(96, 0.2)


def c2_input_density(g):
    """A neon-like closed-shell spin density (5 electrons per channel): a 1s-like core and an
    n = 2-like shell, normalised on the grid."""
    xs, ys, zs = g.coords()
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).ravel()
    n = np.exp(-2.0 * 9.6 * r) * 2.0 + (r * r) * np.exp(-2.0 * 2.8 * r) * 8.0
    n *= 5.0 / (n.sum() * g.dx * g.dy * g.dz)
    return n


def sample_index(n, count=4096, seed=96):
    return np.sort(np.random.default_rng(seed).choice(n, size=count, replace=False))


def main():
    R = Reference()
    g = c2_grid()
    ra = c2_input_density(g)
    t = time.time()
    r = R.scf_step(g, NUCLEI, FUNCTIONAL, 1, ra, ra)
    print("reference step", round(time.time() - t, 1), "s", r["energy"], r["eigenvalues"])
    idx = sample_index(g.size)
    out = dict(grid=np.array(g.n + g.bounds, dtype=float), nuclei=np.array(NUCLEI), functional=np.array(FUNCTIONAL),
               input_sum=np.array(ra.sum()), idx=idx, energy=r["energy"], eigenvalues=r["eigenvalues"],
               dRho=np.array(r["dRho"]), ks=np.array(r["ks"]))
    for k in ("vh", "vxa", "vxb", "ora", "orb", "ovh"):
        f = r[k]
        out[f"{k}_at"] = f[idx]
        out[f"{k}_sum"] = np.array(f.sum())
        out[f"{k}_sum2"] = np.array((f * f).sum())
    path = os.path.join(HERE, "c2_step.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
